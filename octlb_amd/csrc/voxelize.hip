// MI355X (gfx950) voxelizer: triangle soup -> MaterialKind of every cell of
// a forest, the device form of voxelize_part (part_voxelizer.cpp:71-114)
// over CgalSurfaceMesh (cgal_surface_mesh.cpp).
//
// Per cell (box = CellBounds of the block, centre = mid-point):
//   intersects = some triangle of the part (degenerate ones included, as
//                intersects_box does) meets the closed cell box;
//   inside(m)  = ray parity of mesh m at the centre: along +x, +y and +z,
//                count the non-degenerate triangles the closed ray meets
//                (shared edges and vertices count once per triangle, as
//                RayParityInsideDirectionAabb does), inside when at least
//                two of the three counts are odd;
//   external obstacle: boundary if intersects, else solid if inside, else
//                fluid (MapExternal); internal channel: boundary if
//                intersects, fluid in the lumen (inner mesh), solid in the
//                wall (outer mesh minus lumen), else fluid
//                (MapInternalChannel).
// The reference evaluates these with exact predicates (CGAL's
// Exact_predicates_inexact_constructions_kernel), so the tests are written
// as signs of determinants of the input doubles (exact.h): the box test as
// separating axes (box axes, triangle normal, the nine edge x axis
// products), the ray test as the projected-triangle orientations plus the
// side of the plane, with the ray-in-plane case solved in that plane.
//
// Passes: (1) per triangle its box, exact normal signs and degeneracy;
// (2) per block the triangles whose box meets the block (box list) and, per
// axis, the non-degenerate triangles a ray from inside the block can meet
// (counted, scanned on the host, emitted); (3) one workgroup per block
// classifies its cells with the floating-point filters alone and queues
// the cells where a filter could not decide; (4) one wave per queued cell
// repeats the classification with exact fallbacks, its lanes splitting the
// candidate lists.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "exact.h"
#include "octlb_gpu.h"
#include "octlb_internal.h"

namespace octlb {
namespace {

using exact::kUnsure;
using exact::Sign2;
using exact::Sign3;

#define VOX_CHECK(expr)                                                                       \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess)                                                                     \
      throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(_e) +           \
                               " (voxelize.hip:" + std::to_string(__LINE__) + ")");          \
  } while (0)

constexpr uint8_t kFluid = 0, kSolid = 1, kBoundary = 2;
constexpr int kVoxThreads = 256;

// One triangle of a part, with what every test needs of it.
struct VTri {
  double v[3][3];
  double lo[3], hi[3];
  int nsign[3];  // exact signs of the normal (B - A) x (C - A)
  int degenerate;
  int mesh;      // 0 outer, 1 inner (lumen) surface
};

// ---------------------------------------------------------------- ray test
// Does the closed ray p + t e_a (t >= 0) meet the closed non-degenerate
// triangle?  1, 0, or kUnsure (filters only).
template <bool kExact>
__host__ __device__ int RayHit(const double* p, int a, const VTri& t) {
  const double* A = t.v[0];
  const double* B = t.v[1];
  const double* C = t.v[2];
  const int u = (a + 1) % 3, w = (a + 2) % 3;
  const int na = t.nsign[a];
  const int s = Sign3<kExact>(B, C, p, A);  // normal . (p - A)
  if (s == kUnsure) return kUnsure;
  if (na != 0) {
    // the plane is met at t = -s / n_a: behind the origin when s and n_a agree
    if (s != 0 && s == na) return 0;
    // p's projection along a in the closed projected triangle
    const int e[3] = {Sign2<kExact>(B[u], B[w], A[u], A[w], p[u], p[w], A[u], A[w]),
                      Sign2<kExact>(C[u], C[w], B[u], B[w], p[u], p[w], B[u], B[w]),
                      Sign2<kExact>(A[u], A[w], C[u], C[w], p[u], p[w], C[u], C[w])};
    bool unsure = false;
    for (int k = 0; k < 3; ++k) {
      if (e[k] == -na) return 0;
      if (e[k] == kUnsure) unsure = true;
    }
    return unsure ? kUnsure : 1;
  }
  if (s != 0) return 0;  // parallel plane off the ray
  // the ray lies in the triangle's plane: solve in the coordinate plane
  // (a, q) onto which the triangle projects without collapsing
  const int q = t.nsign[w] != 0 ? u : w;
  bool unsure = false;
  for (int k = 0; k < 3; ++k) {
    const double* X = t.v[k];
    const double* Y = t.v[(k + 1) % 3];
    const double pq = p[q];
    if ((X[q] > pq && Y[q] > pq) || (X[q] < pq && Y[q] < pq)) continue;
    if (X[q] == pq && Y[q] == pq) {
      if (fmax(X[a], Y[a]) >= p[a]) return 1;
      continue;
    }
    // the edge crosses the line through p at a*: (a* - p_a)(Y_q - X_q) has
    // the sign of (X_a - p_a)(Y_q - X_q) - (X_q - p_q)(Y_a - X_a)
    const int num = Sign2<kExact>(X[a], X[q], p[a], pq, Y[a], Y[q], X[a], X[q]);
    if (num == kUnsure) {
      unsure = true;
      continue;
    }
    const int dq = Y[q] > X[q] ? 1 : -1;
    if (num == 0 || num == dq) return 1;
  }
  return unsure ? kUnsure : 0;
}

// ---------------------------------------------------------------- box test
// Does the closed triangle (degenerate ones too) meet the closed box?
template <bool kExact>
__host__ __device__ int BoxHit(const double* lo, const double* hi, const VTri& t) {
  for (int k = 0; k < 3; ++k)
    if (t.hi[k] < lo[k] || t.lo[k] > hi[k]) return 0;
  const double* A = t.v[0];
  const double* B = t.v[1];
  const double* C = t.v[2];
  bool unsure = false;
  if (t.nsign[0] != 0 || t.nsign[1] != 0 || t.nsign[2] != 0) {
    // triangle plane: the box corners nearest and farthest along the normal
    double cmin[3], cmax[3];
    for (int k = 0; k < 3; ++k) {
      cmin[k] = t.nsign[k] > 0 ? lo[k] : hi[k];
      cmax[k] = t.nsign[k] > 0 ? hi[k] : lo[k];
    }
    const int s1 = Sign3<kExact>(B, C, cmin, A);
    if (s1 == 1) return 0;
    const int s2 = Sign3<kExact>(B, C, cmax, A);
    if (s2 == -1) return 0;
    if (s1 == kUnsure || s2 == kUnsure) unsure = true;
  }
  // axes e_k x (Y - X) for the three edges
  for (int e = 0; e < 3; ++e) {
    const double* X = t.v[e];
    const double* Y = t.v[(e + 1) % 3];
    const double* Z = t.v[(e + 2) % 3];
    for (int k = 0; k < 3; ++k) {
      const int i1 = (k + 1) % 3, i2 = (k + 2) % 3;
      if (Y[i1] == X[i1] && Y[i2] == X[i2]) continue;  // edge along e_k: no axis
      // L = (L_i1, L_i2) = (-(Y - X)_i2, (Y - X)_i1); the box's extreme corners along L
      const int l1 = Y[i2] > X[i2] ? -1 : (Y[i2] < X[i2] ? 1 : 0);
      const int l2 = Y[i1] > X[i1] ? 1 : (Y[i1] < X[i1] ? -1 : 0);
      const double mn1 = l1 > 0 ? lo[i1] : hi[i1], mx1 = l1 > 0 ? hi[i1] : lo[i1];
      const double mn2 = l2 > 0 ? lo[i2] : hi[i2], mx2 = l2 > 0 ? hi[i2] : lo[i2];
      // L . (V - c) = (Y - X)_i1 (V - c)_i2 - (Y - X)_i2 (V - c)_i1
      const int xa = Sign2<kExact>(Y[i1], Y[i2], X[i1], X[i2], X[i1], X[i2], mn1, mn2);
      const int za = Sign2<kExact>(Y[i1], Y[i2], X[i1], X[i2], Z[i1], Z[i2], mn1, mn2);
      if (xa == -1 && za == -1) return 0;  // triangle entirely below the box
      const int xb = Sign2<kExact>(Y[i1], Y[i2], X[i1], X[i2], X[i1], X[i2], mx1, mx2);
      const int zb = Sign2<kExact>(Y[i1], Y[i2], X[i1], X[i2], Z[i1], Z[i2], mx1, mx2);
      if (xb == 1 && zb == 1) return 0;  // entirely above
      if ((xa == kUnsure || za == kUnsure) && xa != 0 && xa != 1 && za != 0 && za != 1) unsure = true;
      if ((xb == kUnsure || zb == kUnsure) && xb != 0 && xb != -1 && zb != 0 && zb != -1) unsure = true;
    }
  }
  return unsure ? kUnsure : 1;
}

// ------------------------------------------------------------ per-triangle
__global__ void KTriPrep(const double* tri, int64_t ntri, int64_t nouter, VTri* out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= ntri) return;
  VTri t;
  for (int k = 0; k < 3; ++k)
    for (int d = 0; d < 3; ++d) t.v[k][d] = tri[i * 9 + k * 3 + d];
  for (int d = 0; d < 3; ++d) {
    t.lo[d] = fmin(fmin(t.v[0][d], t.v[1][d]), t.v[2][d]);
    t.hi[d] = fmax(fmax(t.v[0][d], t.v[1][d]), t.v[2][d]);
  }
  for (int a = 0; a < 3; ++a) {
    const int u = (a + 1) % 3, w = (a + 2) % 3;
    t.nsign[a] = Sign2<true>(t.v[1][u], t.v[1][w], t.v[0][u], t.v[0][w], t.v[2][u], t.v[2][w], t.v[0][u],
                             t.v[0][w]);
  }
  t.degenerate = (t.nsign[0] == 0 && t.nsign[1] == 0 && t.nsign[2] == 0) ? 1 : 0;
  t.mesh = i < nouter ? 0 : 1;
  out[i] = t;
}

// ------------------------------------------------------------ candidates
// List l of block b: 0 = box list, 1 + a = ray list along axis a.
__device__ inline bool InList(const VTri& t, const double* bb, int l) {
  if (l == 0) {
    for (int k = 0; k < 3; ++k)
      if (t.hi[k] < bb[k] || t.lo[k] > bb[3 + k]) return false;
    return true;
  }
  if (t.degenerate) return false;
  const int a = l - 1, u = (a + 1) % 3, w = (a + 2) % 3;
  return !(t.hi[u] < bb[u] || t.lo[u] > bb[3 + u] || t.hi[w] < bb[w] || t.lo[w] > bb[3 + w] ||
           t.hi[a] < bb[a]);
}

// EMIT == false: counts[b * 4 + l]; EMIT == true: the indices at offs[b * 4 + l]
template <bool EMIT>
__global__ __launch_bounds__(kVoxThreads) void KLists(const VTri* tris, int64_t ntri, const double* bounds,
                                                      int64_t* counts, const int64_t* offs, int* idx) {
  __shared__ unsigned int cnt[4];
  const int b = blockIdx.x;
  if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
  __syncthreads();
  double bb[6];
  for (int k = 0; k < 6; ++k) bb[k] = bounds[static_cast<size_t>(b) * 6 + k];
  for (int64_t i = threadIdx.x; i < ntri; i += blockDim.x) {
    const VTri& t = tris[i];
    for (int l = 0; l < 4; ++l) {
      if (!InList(t, bb, l)) continue;
      const unsigned int at = atomicAdd(&cnt[l], 1u);
      if (EMIT) idx[offs[b * 4 + l] + at] = static_cast<int>(i);
    }
  }
  __syncthreads();
  if (!EMIT && threadIdx.x < 4) counts[b * 4 + threadIdx.x] = cnt[threadIdx.x];
}

// ------------------------------------------------------------ classification
struct VoxArgs {
  const VTri* tris;
  const double* bounds;
  const int64_t* offs;  // [block][5]: box list, ray lists x, y, z, end
  const int* idx;
  int n;
  int channel;          // internal channel (else external obstacle)
  uint8_t* material;
  int* queue;           // cells the filters left open: block * n^3 + cell
  unsigned int* queued;
};

// CellBounds (part_voxelizer.cpp:33-47) and the centre, in the reference's
// operation order.
__host__ __device__ inline void CellBox(const double* ob, int n, int i, int j, int k, double* lo, double* hi,
                                        double* c) {
  const int ijk[3] = {i, j, k};
  for (int d = 0; d < 3; ++d) {
    const double dx = (ob[3 + d] - ob[d]) / static_cast<double>(n);
    lo[d] = ob[d] + static_cast<double>(ijk[d]) * dx;
    hi[d] = lo[d] + dx;
    c[d] = 0.5 * (lo[d] + hi[d]);
  }
}

// A ray from p along +a can only meet a triangle whose box holds p's
// projection and reaches p along a (exact comparisons).
__device__ inline bool RayMayHit(const double* p, int a, const VTri& t) {
  const int u = (a + 1) % 3, w = (a + 2) % 3;
  return !(t.hi[a] < p[a] || t.lo[u] > p[u] || t.hi[u] < p[u] || t.lo[w] > p[w] || t.hi[w] < p[w]);
}

// MaterialKind of one cell, or -1 when a filter could not decide.
template <bool kExact>
__device__ int Classify(const VoxArgs& a, int b, int cell) {
  const int n = a.n;
  const int i = cell % n, j = (cell / n) % n, k = cell / (n * n);
  double lo[3], hi[3], c[3];
  CellBox(a.bounds + static_cast<size_t>(b) * 6, n, i, j, k, lo, hi, c);
  const int64_t* o = a.offs + static_cast<size_t>(b) * 5;
  bool open = false;
  for (int64_t e = o[0]; e < o[1]; ++e) {
    const int h = BoxHit<kExact>(lo, hi, a.tris[a.idx[e]]);
    if (h == 1) return kBoundary;
    if (h == kUnsure) open = true;
  }
  if (open) return -1;
  // ray parity per mesh and axis
  int odd[2][3], unsure[2][3];
  for (int m = 0; m < 2; ++m)
    for (int ax = 0; ax < 3; ++ax) odd[m][ax] = unsure[m][ax] = 0;
  for (int ax = 0; ax < 3; ++ax)
    for (int64_t e = o[1 + ax]; e < o[2 + ax]; ++e) {
      const VTri& t = a.tris[a.idx[e]];
      if (!RayMayHit(c, ax, t)) continue;
      const int h = RayHit<kExact>(c, ax, t);
      if (h == 1) odd[t.mesh][ax] ^= 1;
      if (h == kUnsure) unsure[t.mesh][ax] = 1;
    }
  // inside votes: decided when the certain axes settle the majority
  int inside[2];
  for (int m = 0; m < 2; ++m) {
    int yes = 0, no = 0;
    for (int ax = 0; ax < 3; ++ax) {
      if (unsure[m][ax]) continue;
      if (odd[m][ax]) ++yes; else ++no;
    }
    inside[m] = yes >= 2 ? 1 : (no >= 2 ? 0 : -1);
  }
  if (!a.channel) {
    if (inside[0] < 0) return -1;
    return inside[0] ? kSolid : kFluid;
  }
  if (inside[1] < 0) return -1;
  if (inside[1]) return kFluid;  // lumen
  if (inside[0] < 0) return -1;
  return inside[0] ? kSolid : kFluid;
}

__global__ __launch_bounds__(kVoxThreads) void KClassify(VoxArgs a) {
  const int b = blockIdx.x;
  const int cells = a.n * a.n * a.n;
  for (int cell = threadIdx.x; cell < cells; cell += blockDim.x) {
    const int m = Classify<false>(a, b, cell);
    const size_t at = static_cast<size_t>(b) * cells + cell;
    if (m >= 0) {
      a.material[at] = static_cast<uint8_t>(m);
    } else {
      a.queue[atomicAdd(a.queued, 1u)] = static_cast<int>(at);
    }
  }
}

// One wave per queued cell: the lanes split the candidate lists and combine
// any-hit and hit parities with ballots; every predicate the filter leaves
// open is decided exactly.
__global__ __launch_bounds__(64) void KClassifyExact(VoxArgs a) {
  const int lane = threadIdx.x;
  const int at = a.queue[blockIdx.x];
  const int n = a.n, cells = n * n * n;
  const int b = at / cells, cell = at % cells;
  double lo[3], hi[3], c[3];
  CellBox(a.bounds + static_cast<size_t>(b) * 6, n, cell % n, (cell / n) % n, cell / (n * n), lo, hi, c);
  const int64_t* o = a.offs + static_cast<size_t>(b) * 5;
  bool hit = false;
  for (int64_t e = o[0] + lane; e < o[1]; e += 64) hit = hit || BoxHit<true>(lo, hi, a.tris[a.idx[e]]) == 1;
  if (__any(hit)) {
    if (lane == 0) a.material[at] = kBoundary;
    return;
  }
  int inside[2];
  for (int m = 0; m < 2; ++m) {
    int votes = 0;
    for (int ax = 0; ax < 3; ++ax) {
      int odd = 0;
      for (int64_t e = o[1 + ax] + lane; e < o[2 + ax]; e += 64) {
        const VTri& t = a.tris[a.idx[e]];
        if (t.mesh == m && RayMayHit(c, ax, t) && RayHit<true>(c, ax, t) == 1) odd ^= 1;
      }
      votes += static_cast<int>(__popcll(__ballot(odd)) & 1);
    }
    inside[m] = votes >= 2;
  }
  if (lane == 0)
    a.material[at] = a.channel ? (inside[1] ? kFluid : (inside[0] ? kSolid : kFluid))
                               : (inside[0] ? kSolid : kFluid);
}

// outside_sample_fraction (cgal_surface_mesh.cpp): s^3 points over the
// outer soup's box, inside by ray parity against every outer triangle.
__global__ void KSampleInside(const VTri* tris, int64_t nouter, const double* box, int s, int* inside) {
  const int q = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (q >= s * s * s) return;
  const int ix = q % s, iy = (q / s) % s, iz = q / (s * s);
  const double p[3] = {box[0] + static_cast<double>(ix) * ((box[3] - box[0]) / static_cast<double>(s - 1)),
                       box[1] + static_cast<double>(iy) * ((box[4] - box[1]) / static_cast<double>(s - 1)),
                       box[2] + static_cast<double>(iz) * ((box[5] - box[2]) / static_cast<double>(s - 1))};
  int votes = 0;
  for (int ax = 0; ax < 3; ++ax) {
    int odd = 0;
    for (int64_t e = 0; e < nouter; ++e)
      if (!tris[e].degenerate && RayHit<true>(p, ax, tris[e]) == 1) odd ^= 1;
    votes += odd;
  }
  inside[q] = votes >= 2 ? 1 : 0;
}

template <typename P>
P* VoxAlloc(size_t count) {
  void* p = nullptr;
  VOX_CHECK(hipMalloc(&p, (count ? count : 1) * sizeof(P)));
  return static_cast<P*>(p);
}

struct VoxBuffers {
  std::vector<void*> ptrs;
  template <typename P>
  P* Get(size_t count) {
    P* p = VoxAlloc<P>(count);
    ptrs.push_back(p);
    return p;
  }
  ~VoxBuffers() {
    for (void* p : ptrs) (void)hipFree(p);
  }
};

constexpr double kMinLumenOutsideFraction = 0.05;  // part_voxelizer.cpp:10-11
constexpr int kWatertightSampleResolution = 5;

}  // namespace

int VoxelizePart(int32_t num_blocks, int32_t n, const double* bounds, int64_t num_triangles,
                 const double* triangles, int64_t num_inner, const double* inner, int32_t role, uint8_t* material,
                 int64_t* num_exact) {
  if (num_blocks <= 0 || n <= 0 || !bounds || !material) throw std::invalid_argument("bad forest arguments");
  if (role != OCTLB_PART_INTERNAL_CHANNEL && role != OCTLB_PART_EXTERNAL_OBSTACLE)
    throw std::invalid_argument("unknown part role");
  // validate_geometry_part (part_voxelizer.cpp:51-69)
  if (num_triangles <= 0 || !triangles) throw std::invalid_argument("geometry part has no triangles");
  const bool channel = role == OCTLB_PART_INTERNAL_CHANNEL;
  if (channel && (num_inner <= 0 || !inner)) throw std::invalid_argument("kInternalChannel requires inner_cavity_soup");
  if (!channel) num_inner = 0;
  const int64_t ntri = num_triangles + num_inner;
  if (ntri > static_cast<int64_t>(INT32_MAX)) throw std::invalid_argument("too many triangles (> 2^31 - 1)");
  const size_t cells = static_cast<size_t>(n) * n * n;
  const size_t total = static_cast<size_t>(num_blocks) * cells;
  if (total > static_cast<size_t>(INT32_MAX)) throw std::invalid_argument("forest too large (cells > 2^31)");

  VoxBuffers buf;
  double* d_raw = buf.Get<double>(static_cast<size_t>(ntri) * 9);
  VOX_CHECK(hipMemcpy(d_raw, triangles, static_cast<size_t>(num_triangles) * 9 * sizeof(double),
                      hipMemcpyHostToDevice));
  if (num_inner > 0)
    VOX_CHECK(hipMemcpy(d_raw + num_triangles * 9, inner, static_cast<size_t>(num_inner) * 9 * sizeof(double),
                        hipMemcpyHostToDevice));
  VTri* d_tri = buf.Get<VTri>(static_cast<size_t>(ntri));
  hipLaunchKernelGGL(KTriPrep, dim3(static_cast<unsigned>((ntri + 255) / 256)), dim3(256), 0, nullptr, d_raw, ntri,
                     num_triangles, d_tri);
  VOX_CHECK(hipGetLastError());

  if (channel) {
    // the outer mesh must leave lumen/exterior volume (outside fraction)
    double box[6] = {triangles[0], triangles[1], triangles[2], triangles[0], triangles[1], triangles[2]};
    for (int64_t t = 0; t < num_triangles * 3; ++t)
      for (int d = 0; d < 3; ++d) {
        box[d] = std::min(box[d], triangles[t * 3 + d]);
        box[3 + d] = std::max(box[3 + d], triangles[t * 3 + d]);
      }
    const int s = kWatertightSampleResolution, pts = s * s * s;
    double* d_box = buf.Get<double>(6);
    int* d_in = buf.Get<int>(pts);
    VOX_CHECK(hipMemcpy(d_box, box, sizeof(box), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(KSampleInside, dim3((pts + 63) / 64), dim3(64), 0, nullptr, d_tri, num_triangles, d_box, s,
                       d_in);
    VOX_CHECK(hipGetLastError());
    std::vector<int> in(pts);
    VOX_CHECK(hipMemcpy(in.data(), d_in, pts * sizeof(int), hipMemcpyDeviceToHost));
    int outside = 0;
    for (int v : in) outside += v ? 0 : 1;
    if (static_cast<double>(outside) / static_cast<double>(pts) < kMinLumenOutsideFraction)
      throw std::invalid_argument(
          "kInternalChannel mesh has insufficient exterior/lumen volume (solid-wall channel expected)");
  }

  double* d_bounds = buf.Get<double>(static_cast<size_t>(num_blocks) * 6);
  VOX_CHECK(hipMemcpy(d_bounds, bounds, static_cast<size_t>(num_blocks) * 6 * sizeof(double), hipMemcpyHostToDevice));
  int64_t* d_cnt = buf.Get<int64_t>(static_cast<size_t>(num_blocks) * 4);
  hipLaunchKernelGGL(KLists<false>, dim3(num_blocks), dim3(kVoxThreads), 0, nullptr, d_tri, ntri, d_bounds, d_cnt,
                     nullptr, nullptr);
  VOX_CHECK(hipGetLastError());
  std::vector<int64_t> cnt(static_cast<size_t>(num_blocks) * 4);
  VOX_CHECK(hipMemcpy(cnt.data(), d_cnt, cnt.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
  // offsets: emission slots [block][list], classification ranges [block][5]
  std::vector<int64_t> emit(cnt.size()), ranges(static_cast<size_t>(num_blocks) * 5);
  int64_t at = 0;
  for (int b = 0; b < num_blocks; ++b) {
    for (int l = 0; l < 4; ++l) {
      ranges[static_cast<size_t>(b) * 5 + l] = at;
      emit[static_cast<size_t>(b) * 4 + l] = at;
      at += cnt[static_cast<size_t>(b) * 4 + l];
    }
    ranges[static_cast<size_t>(b) * 5 + 4] = at;
  }
  if (at > INT32_MAX) throw std::invalid_argument("candidate lists too long");
  int64_t* d_emit = buf.Get<int64_t>(emit.size());
  int64_t* d_ranges = buf.Get<int64_t>(ranges.size());
  int* d_idx = buf.Get<int>(static_cast<size_t>(at));
  VOX_CHECK(hipMemcpy(d_emit, emit.data(), emit.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  VOX_CHECK(hipMemcpy(d_ranges, ranges.data(), ranges.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(KLists<true>, dim3(num_blocks), dim3(kVoxThreads), 0, nullptr, d_tri, ntri, d_bounds, nullptr,
                     d_emit, d_idx);
  VOX_CHECK(hipGetLastError());

  VoxArgs a{};
  a.tris = d_tri;
  a.bounds = d_bounds;
  a.offs = d_ranges;
  a.idx = d_idx;
  a.n = n;
  a.channel = channel ? 1 : 0;
  a.material = buf.Get<uint8_t>(total);
  a.queue = buf.Get<int>(total);
  a.queued = buf.Get<unsigned int>(1);
  VOX_CHECK(hipMemset(a.queued, 0, sizeof(unsigned int)));
  hipLaunchKernelGGL(KClassify, dim3(num_blocks), dim3(kVoxThreads), 0, nullptr, a);
  VOX_CHECK(hipGetLastError());
  unsigned int queued = 0;
  VOX_CHECK(hipMemcpy(&queued, a.queued, sizeof(queued), hipMemcpyDeviceToHost));
  if (queued > 0) {
    hipLaunchKernelGGL(KClassifyExact, dim3(queued), dim3(64), 0, nullptr, a);
    VOX_CHECK(hipGetLastError());
  }
  VOX_CHECK(hipMemcpy(material, a.material, total, hipMemcpyDeviceToHost));
  if (num_exact) *num_exact = queued;
  return OCTLB_OK;
}

}  // namespace octlb

extern "C" int octlb_voxelize_part(int32_t num_blocks, int32_t n, const double* bounds, int64_t num_triangles,
                                   const double* triangles, int64_t num_inner_triangles,
                                   const double* inner_triangles, int32_t role, uint8_t* material,
                                   int64_t* num_exact) {
  try {
    return octlb::VoxelizePart(num_blocks, n, bounds, num_triangles, triangles, num_inner_triangles,
                               inner_triangles, role, material, num_exact);
  } catch (const std::invalid_argument& e) {
    return octlb::SetError(OCTLB_EINVAL, e.what());
  } catch (const std::exception& e) {
    return octlb::SetError(OCTLB_EHIP, e.what());
  }
}
