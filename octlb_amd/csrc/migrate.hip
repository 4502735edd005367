// MI355X (gfx950) state migration after a forest change: every block of the
// adapted forest takes its populations from the old forest's device buffers,
// one workgroup per new block and one launch per rule (copy, refine,
// coarsen).  The host only matches blocks (level + integer origin, hashed)
// and uploads one task per new block with the block lists of the rules.
//
// The rules (no reference counterpart; the reference plans its dynamic-AMR
// rebuild in doc/prd/octlb-framework.md:180,521):
//   copy     same level and bounds: the halo-inclusive view as stored
//   refine   every slot of the fine block, halo included, from the parent
//            cell under it: feq(rho, u) + s_p fneq with s_p = (tau - 1/4)/tau,
//            the coupler's prolongation rule (level_coupler.cpp:163-214)
//   coarsen  each coarse cell from its eight fine cells: mean rho, the
//            momentum-conserving mean u and s_r times the mean fneq
//            (s_r = 1/s_p); halo slots repeat the nearest core cell
// A view slot of an old block is what the step would read there: its core,
// the exchanged neighbour's boundary layer, else the block's shell, else 0.
// Arithmetic is FP64 in both precisions, in the same order as the rules'
// CPU statement, so the results are bit for bit those of that statement.
#include "lbm_device.h"

namespace octlb {

namespace {

// Padded 10^3 slot -> shell (apron) index, -1 for the 8^3 core.
struct ApronInverse {
  short ap[kBoxCells];
};
constexpr ApronInverse MakeApronInverse() {
  ApronInverse inv{};
  for (int s = 0; s < kBoxCells; ++s) inv.ap[s] = -1;
  const ApronTable t = MakeApron();
  for (int j = 0; j < kApron; ++j) inv.ap[t.idx[j]] = static_cast<short>(j);
  return inv;
}
__constant__ ApronInverse kApronInv = MakeApronInverse();
__constant__ ApronTable kApronFwd = MakeApron();

// Block whose core supplies halo cell (x,y,z) of a blocked-mode block through
// a face or edge exchange (BlockTab::nbr), or -1.
__device__ inline int ExchangedFrom(const BlockTab& bt, int x, int y, int z) {
  constexpr int n = kTile;
  const int ox = (x < 0 || x >= n), oy = (y < 0 || y >= n), oz = (z < 0 || z >= n);
  const int nout = ox + oy + oz;
  if (nout == 1) {
    const int axis = ox ? 0 : (oy ? 1 : 2);
    const int v = axis == 0 ? x : (axis == 1 ? y : z);
    return bt.nbr[axis * 2 + (v >= n ? 1 : 0)];
  }
  if (nout == 2) {
    const int a1 = ox ? 0 : 1, a2 = oz ? 2 : 1;
    const int v1 = a1 == 0 ? x : y, v2 = a2 == 2 ? z : y;
    const int pair = (a1 == 0) ? (a2 == 1 ? 0 : 1) : 2;
    return bt.nbr[6 + pair * 4 + (v1 >= n ? 1 : 0) * 2 + (v2 >= n ? 1 : 0)];
  }
  return -1;
}

// Slot (x,y,z) of old block b's halo-inclusive view, x,y,z in [-h, n+h).
template <typename T>
__device__ void ReadView(const MigArgs<T>& a, int b, int x, int y, int z, double* f) {
  const MigSide<T>& s = a.src;
  const BlockTab& bt = s.bt[b];
  const T* pops = s.pops[s.par.v[bt.level]];
  const size_t plane = static_cast<size_t>(a.plane);
  if (a.padded) {
    const size_t base = static_cast<size_t>(b) * kQ * plane + ((x + a.h) * a.NP + y + a.h) * a.NP + z + a.h;
    for (int i = 0; i < kQ; ++i) f[i] = static_cast<double>(pops[base + i * plane]);
    return;
  }
  constexpr int n = kTile;
  if (x >= 0 && x < n && y >= 0 && y < n && z >= 0 && z < n) {
    const size_t base = static_cast<size_t>(b) * kQ * plane + (x * n + y) * n + z;
    for (int i = 0; i < kQ; ++i) f[i] = static_cast<double>(pops[base + i * plane]);
    return;
  }
  const int nb = ExchangedFrom(bt, x, y, z);
  if (nb >= 0) {
    const T* np = s.pops[s.par.v[s.bt[nb].level]];
    const int wx = x < 0 ? n - 1 : (x >= n ? 0 : x);
    const int wy = y < 0 ? n - 1 : (y >= n ? 0 : y);
    const int wz = z < 0 ? n - 1 : (z >= n ? 0 : z);
    const size_t base = static_cast<size_t>(nb) * kQ * plane + (wx * n + wy) * n + wz;
    for (int i = 0; i < kQ; ++i) f[i] = static_cast<double>(np[base + i * plane]);
    return;
  }
  if (bt.shell >= 0) {
    const T* sh = s.shells[s.par.v[bt.level]];
    const int ap = kApronInv.ap[((x + 1) * kBox + y + 1) * kBox + z + 1];
    const size_t base = static_cast<size_t>(bt.shell) * kQ * kApron + ap;
    for (int i = 0; i < kQ; ++i) f[i] = static_cast<double>(sh[base + static_cast<size_t>(i) * kApron]);
    return;
  }
  for (int i = 0; i < kQ; ++i) f[i] = 0.0;
}

// Stores slot (x,y,z) of new block b where the layout keeps it: every slot
// (padded), core cells, and shell slots of blocks with a shell.
template <typename T>
__device__ void WriteView(const MigArgs<T>& a, int b, int x, int y, int z, const double* f) {
  const MigSide<T>& d = a.dst;
  const BlockTab& bt = d.bt[b];
  T* pops = d.pops[d.par.v[bt.level]];
  const size_t plane = static_cast<size_t>(a.plane);
  if (a.padded) {
    const size_t base = static_cast<size_t>(b) * kQ * plane + ((x + a.h) * a.NP + y + a.h) * a.NP + z + a.h;
    for (int i = 0; i < kQ; ++i) pops[base + i * plane] = static_cast<T>(f[i]);
    return;
  }
  constexpr int n = kTile;
  if (x >= 0 && x < n && y >= 0 && y < n && z >= 0 && z < n) {
    const size_t base = static_cast<size_t>(b) * kQ * plane + (x * n + y) * n + z;
    for (int i = 0; i < kQ; ++i) pops[base + i * plane] = static_cast<T>(f[i]);
    return;
  }
  T* sh = d.shells[d.par.v[bt.level]];
  const int ap = kApronInv.ap[((x + 1) * kBox + y + 1) * kBox + z + 1];
  const size_t base = static_cast<size_t>(bt.shell) * kQ * kApron + ap;
  for (int i = 0; i < kQ; ++i) sh[base + static_cast<size_t>(i) * kApron] = static_cast<T>(f[i]);
}

// olb::lbm::computeRhoU of shifted populations and the non-equilibrium part.
__device__ inline void Macro(const double* f, double& rho, double* u, double* fneq) {
  rho = 0.0;
  for (int i = 0; i < kQ; ++i) rho += f[i];
  rho += 1.0;
  u[0] = u[1] = u[2] = 0.0;
  for (int i = 0; i < kQ; ++i)
    for (int d = 0; d < 3; ++d) u[d] += f[i] * Ci(i, d);
  for (int d = 0; d < 3; ++d) u[d] /= rho;
  const double usqr = NormSqr(u);
  for (int i = 0; i < kQ; ++i) fneq[i] = f[i] - Feq<double>(i, rho, u, usqr);
}

__device__ inline int FloorHalf(int g) { return g >= 0 ? g / 2 : -((1 - g) / 2); }

// New block's slot (x,y,z) from the old forest, by the task's rule.
template <typename T, int Mode>
__device__ void MigrateSlot(const MigArgs<T>& a, const MigTask& t, int b, int x, int y, int z) {
  const int n = a.n;
  double out[kQ];
  if constexpr (Mode == kMigCopy) {
    ReadView(a, t.src[0], x, y, z, out);
  } else if constexpr (Mode == kMigRefine) {
    // fine index g = oct n + x in the parent's extent; parent cell g / 2
    double in[kQ], rho, u[3], fneq[kQ];
    ReadView(a, t.src[0], FloorHalf(t.oct[0] * n + x), FloorHalf(t.oct[1] * n + y), FloorHalf(t.oct[2] * n + z), in);
    Macro(in, rho, u, fneq);
    const double usqr = NormSqr(u);
    for (int i = 0; i < kQ; ++i) out[i] = Feq<double>(i, rho, u, usqr) + a.sp * fneq[i];
  } else {
    // blocked layout only (n = 8): the eight fine cells under a coarse cell
    // lie in one child; halo slots repeat the nearest core cell
    constexpr int m = kTile;
    const int cx = min(max(x, 0), m - 1), cy = min(max(y, 0), m - 1), cz = min(max(z, 0), m - 1);
    const int kid = t.src[(2 * cx) / m + 2 * ((2 * cy) / m) + 4 * ((2 * cz) / m)];
    double acc[4 + kQ];
    for (int k = 0; k < 4 + kQ; ++k) acc[k] = 0.0;
    // the child's cells in its raster order
    for (int dx = 0; dx < 2; ++dx)
      for (int dy = 0; dy < 2; ++dy)
        for (int dz = 0; dz < 2; ++dz) {
          double in[kQ], rho, u[3], fneq[kQ];
          ReadView(a, kid, (2 * cx + dx) % m, (2 * cy + dy) % m, (2 * cz + dz) % m, in);
          Macro(in, rho, u, fneq);
          acc[0] += rho;
          for (int d = 0; d < 3; ++d) acc[1 + d] += rho * u[d];
          for (int i = 0; i < kQ; ++i) acc[4 + i] += fneq[i];
        }
    const double rho = acc[0] / 8.0;
    const double u[3] = {acc[1] / acc[0], acc[2] / acc[0], acc[3] / acc[0]};
    const double usqr = NormSqr(u);
    for (int i = 0; i < kQ; ++i) out[i] = Feq<double>(i, rho, u, usqr) + a.sr * (acc[4 + i] / 8.0);
  }
  WriteView(a, b, x, y, z, out);
}

// Halo slots of a blocked-mode block with a shell, in shell order.
template <typename T, int Mode>
__device__ void MigrateShell(const MigArgs<T>& a, const MigTask& t, int b) {
  if (a.dst.bt[b].shell < 0) return;
  for (int ap = threadIdx.x; ap < kApron; ap += blockDim.x) {
    const int l = kApronFwd.idx[ap];
    MigrateSlot<T, Mode>(a, t, b, l / (kBox * kBox) - 1, (l / kBox) % kBox - 1, l % kBox - 1);
  }
}

// Blocks with the same level and bounds: the block's storage (its whole
// view in the padded layout, the 19 x 512 core planes otherwise) is one
// contiguous run, copied 16 bytes per lane; a blocked-mode shell takes the
// old block's view slots.
template <typename T>
__global__ __launch_bounds__(256) void KMigrateCopy(MigArgs<T> a, const int* list) {
  const int b = list[blockIdx.x];
  const MigTask t = a.task[b];
  const int s = t.src[0];
  const size_t elems = static_cast<size_t>(kQ) * a.plane;
  const T* from = a.src.pops[a.src.par.v[a.src.bt[s].level]] + static_cast<size_t>(s) * elems;
  T* to = a.dst.pops[a.dst.par.v[a.dst.bt[b].level]] + static_cast<size_t>(b) * elems;
  if ((elems * sizeof(T)) % sizeof(uint4) == 0) {
    // every block starts 16-byte aligned (the buffers are)
    const size_t words = elems * sizeof(T) / sizeof(uint4);
    for (size_t w = threadIdx.x; w < words; w += blockDim.x)
      reinterpret_cast<uint4*>(to)[w] = reinterpret_cast<const uint4*>(from)[w];
  } else {
    for (size_t e = threadIdx.x; e < elems; e += blockDim.x) to[e] = from[e];
  }
  if (!a.padded) MigrateShell<T, kMigCopy>(a, t, b);
}

// Refined or coarsened blocks: one lane per stored slot, core cells in
// storage order.
template <typename T, int Mode>
__global__ __launch_bounds__(512) void KMigrateCells(MigArgs<T> a, const int* list) {
  const int b = list[blockIdx.x];
  const MigTask t = a.task[b];
  if (a.padded) {
    const int E = a.n + 2 * a.h;
    for (int s = threadIdx.x; s < E * E * E; s += blockDim.x)
      MigrateSlot<T, Mode>(a, t, b, s / (E * E) - a.h, (s / E) % E - a.h, s % E - a.h);
    return;
  }
  for (int c = threadIdx.x; c < kTileCells; c += blockDim.x)
    MigrateSlot<T, Mode>(a, t, b, c >> 6, (c >> 3) & 7, c & 7);
  MigrateShell<T, Mode>(a, t, b);
}

}  // namespace

template <typename T>
void LaunchMigrate(const MigArgs<T>& a, const int* list, const int count[3], hipStream_t s) {
  if (count[0] > 0) hipLaunchKernelGGL(KMigrateCopy<T>, dim3(count[0]), dim3(256), 0, s, a, list);
  if (count[1] > 0)
    hipLaunchKernelGGL((KMigrateCells<T, kMigRefine>), dim3(count[1]), dim3(512), 0, s, a, list + count[0]);
  if (count[2] > 0)
    hipLaunchKernelGGL((KMigrateCells<T, kMigCoarsen>), dim3(count[2]), dim3(512), 0, s, a,
                       list + count[0] + count[1]);
}

template void LaunchMigrate<double>(const MigArgs<double>&, const int*, const int[3], hipStream_t);
template void LaunchMigrate<float>(const MigArgs<float>&, const int*, const int[3], hipStream_t);

}  // namespace octlb
