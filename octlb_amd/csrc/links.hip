// Bouzidi wall fractions from a triangle soup on the GPU: the device form of
// BouzidiLinkData::Build (bouzidi_link_data.cpp:139-233).
//
// The reference visits every fluid cell, finds the nearest point of the soup
// (closest point over the three edges of every triangle, first minimum kept)
// and, for each population whose in-block pull-source neighbour x + c_opp(i)
// is solid-like, stores q = ((w - x) . c dx) / |c dx|^2 when it lies in
// (1e-8, 1 - 1e-8).  Only cells with a solid-like in-block neighbour can
// produce a link, so the device pass
//   1. compacts those candidate cells per block with wave ballots and
//      prefix counts (KCandidates, then a single-workgroup scan),
//   2. gives each candidate one wave: lanes stride over the triangles staged
//      in LDS, keep their first minimum, and a butterfly reduction picks the
//      smallest distance, ties to the lowest triangle index -- the
//      sequential first minimum (KNearest); lanes 0..17 then evaluate the
//      18 populations,
//   3. compacts the links in (block, cell, population) order, the dense
//      q_frac order (KEmit).
// Every floating-point expression keeps the reference's operand order and
// the file is built without FMA contraction, so q is bit-identical.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "d3q19.h"
#include "octlb_gpu.h"
#include "octlb_internal.h"

namespace octlb {
namespace {

#define LINK_CHECK(expr)                                                                      \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess)                                                                     \
      throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(_e) +           \
                               " (links.hip:" + std::to_string(__LINE__) + ")");             \
  } while (0)

constexpr int kLinkThreads = 256;      // 4 waves
constexpr int kTriChunk = kLinkThreads;  // triangles staged in LDS per pass

__device__ __forceinline__ bool SolidLike(uint8_t m) { return m == 1 || m == 2; }  // kSolid, kBoundary

// Material of cell (i, j, k) in MaterialField order (material_field.cpp:16-31).
__device__ __forceinline__ uint8_t Mat(const uint8_t* mat, int n, int i, int j, int k) {
  return mat[(static_cast<size_t>(k) * n + j) * n + i];
}

// A fluid cell with a solid-like in-block neighbour at x + c_opp(i) for some
// i (NeighborMaterial: outside the block counts as fluid).
__device__ bool IsCandidate(const uint8_t* mat, int n, int i, int j, int k) {
  if (Mat(mat, n, i, j, k) != 0) return false;
  bool any = false;
#pragma unroll
  for (int ip = 1; ip < kQ; ++ip) {
    const int o = Opp(ip);
    const int ni = i + Cx(o), nj = j + Cy(o), nk = k + Cz(o);
    if (ni < 0 || ni >= n || nj < 0 || nj >= n || nk < 0 || nk >= n) continue;
    any = any || SolidLike(Mat(mat, n, ni, nj, nk));
  }
  return any;
}

// Per block: count (WRITE = false) or write (WRITE = true) the candidate
// cells in increasing cell index (i*n + j)*n + k.
template <bool WRITE>
__global__ __launch_bounds__(kLinkThreads) void KCandidates(int n, const uint8_t* material, int* counts,
                                                             const int64_t* offsets, int2* cand) {
  __shared__ int wave_sum[kLinkThreads / 64];
  const int b = blockIdx.x;
  const int ncell = n * n * n;
  const uint8_t* mat = material + static_cast<size_t>(b) * ncell;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t run = WRITE ? offsets[b] : 0;
  for (int base = 0; base < ncell; base += kLinkThreads) {
    const int c = base + static_cast<int>(threadIdx.x);
    bool flag = false;
    if (c < ncell) flag = IsCandidate(mat, n, c / (n * n), (c / n) % n, c % n);
    const uint64_t m = __ballot(flag);
    if (lane == 0) wave_sum[wave] = __popcll(m);
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < kLinkThreads / 64; ++w) {
      before += w < wave ? wave_sum[w] : 0;
      total += wave_sum[w];
    }
    if (WRITE && flag) {
      const int rank = __popcll(m & ((uint64_t(1) << lane) - 1));
      cand[run + before + rank] = make_int2(b, c);
    }
    run += total;
    __syncthreads();  // wave_sum is rewritten by the next chunk
  }
  if (!WRITE && threadIdx.x == 0) counts[b] = static_cast<int>(run);
}

// Exclusive prefix sum of count ints into int64 offsets, total appended:
// one workgroup, chunks of 1024 in order (deterministic).
__global__ __launch_bounds__(1024) void KScan(const int* in, int64_t count, int64_t* out) {
  __shared__ int64_t part[1024];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < count; base += 1024) {
    const int64_t i = base + threadIdx.x;
    const int64_t v = i < count ? in[i] : 0;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int64_t add = threadIdx.x >= static_cast<unsigned>(off) ? part[threadIdx.x - off] : 0;
      __syncthreads();
      part[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < count) out[i] = carry + part[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[count] = carry;
}

// ClosestPointOnSegment (bouzidi_link_data.cpp:19-40), std::clamp(t, 0, 1).
__device__ __forceinline__ void SegClosest(const double* p, double ax, double ay, double az, double bx, double by,
                                           double bz, double* o) {
  const double abx = bx - ax, aby = by - ay, abz = bz - az;
  const double apx = p[0] - ax, apy = p[1] - ay, apz = p[2] - az;
  const double ab2 = abx * abx + aby * aby + abz * abz;
  double t = 0.0;
  if (ab2 > 0.0) {
    t = (apx * abx + apy * aby + apz * abz) / ab2;
    t = t < 0.0 ? 0.0 : (1.0 < t ? 1.0 : t);
  }
  o[0] = ax + t * abx;
  o[1] = ay + t * aby;
  o[2] = az + t * abz;
}

// ClosestPointOnTriangle (bouzidi_link_data.cpp:42-76): nearest of the three
// edge points in edge order v0v1, v1v2, v2v0, first kept on ties; v0 when no
// distance is below DBL_MAX.
__device__ __forceinline__ void TriClosest(const double* t, const double* p, double* o) {
  double best = DBL_MAX;
  o[0] = t[0];
  o[1] = t[1];
  o[2] = t[2];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int a = 3 * e, b = 3 * ((e + 1) % 3);
    double c[3];
    SegClosest(p, t[a], t[a + 1], t[a + 2], t[b], t[b + 1], t[b + 2], c);
    const double dx = p[0] - c[0], dy = p[1] - c[1], dz = p[2] - c[2];
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < best) {
      best = d2;
      o[0] = c[0];
      o[1] = c[1];
      o[2] = c[2];
    }
  }
}

// One wave per candidate cell: nearest soup point, then the 18 populations.
// Triangles are staged in LDS with a bounding box; a lane skips the
// triangles that the wave's best point so far already beats strictly (the
// result is the same as the full scan: a skipped triangle has no point as
// close as that best, so it can be neither a closer nor an equal minimum).
// qd[cand][i-1] = q or -1; cnt[cand] = number of links of the cell.
__global__ __launch_bounds__(kLinkThreads) void KNearest(const int2* cand, int64_t ncand, int n, const double* bounds,
                                                         const uint8_t* material, const double* tris, int64_t ntri,
                                                         double* qd, int* cnt) {
  __shared__ double st[kTriChunk * 9];
  __shared__ double box[kTriChunk][6];  // bounding box (lo, hi) of each staged triangle
  const int lane = threadIdx.x & 63;
  const int64_t ci = static_cast<int64_t>(blockIdx.x) * (kLinkThreads / 64) + (threadIdx.x >> 6);
  const bool live = ci < ncand;
  int b = 0, c = 0, i = 0, j = 0, k = 0;
  double p[3] = {0.0, 0.0, 0.0};
  double dx = 0.0;
  if (live) {
    const int2 e = cand[ci];
    b = e.x;
    c = e.y;
    i = c / (n * n);
    j = (c / n) % n;
    k = c % n;
    const double* bb = bounds + 6 * static_cast<size_t>(b);
    dx = (bb[3] - bb[0]) / static_cast<double>(n);
    const double dy = (bb[4] - bb[1]) / static_cast<double>(n);
    const double dz = (bb[5] - bb[2]) / static_cast<double>(n);
    // CellCenter (bouzidi_link_data.cpp:85-94)
    p[0] = bb[0] + (static_cast<double>(i) + 0.5) * dx;
    p[1] = bb[1] + (static_cast<double>(j) + 0.5) * dy;
    p[2] = bb[2] + (static_cast<double>(k) + 0.5) * dz;
  }
  double bd = DBL_MAX;
  double pb = DBL_MAX;  // the wave's best distance at the end of the last chunk
  int64_t bt = INT64_MAX;
  double w[3] = {p[0], p[1], p[2]};
  for (int64_t base = 0; base < ntri; base += kTriChunk) {
    const int64_t nt = ntri - base < kTriChunk ? ntri - base : kTriChunk;
    __syncthreads();  // the previous chunk is no longer read
    for (int e = threadIdx.x; e < nt * 9; e += kLinkThreads) st[e] = tris[base * 9 + e];
    if (threadIdx.x < nt) {
      const double* v = tris + (base + threadIdx.x) * 9;
      for (int d = 0; d < 3; ++d) {
        box[threadIdx.x][d] = fmin(v[d], fmin(v[3 + d], v[6 + d]));
        box[threadIdx.x][3 + d] = fmax(v[d], fmax(v[3 + d], v[6 + d]));
      }
    }
    __syncthreads();
    if (live) {
      for (int r = lane; r < nt; r += 64) {
        // Skip a triangle whose bounding box is farther (squared distance,
        // lowered by 1e-12 relative -- far beyond its rounding error) than
        // the wave's best point: every edge point of it is, so it can be
        // neither a closer nor an equal minimum.
        double lb = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double g = fmax(fmax(box[r][d] - p[d], p[d] - box[r][3 + d]), 0.0);
          lb += g * g;
        }
        if (lb * (1.0 - 1.0e-12) > fmin(bd, pb)) continue;
        double q[3];
        TriClosest(&st[r * 9], p, q);
        const double d2 = (p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) +
                          (p[2] - q[2]) * (p[2] - q[2]);
        if (d2 < bd) {
          bd = d2;
          bt = base + r;
          w[0] = q[0];
          w[1] = q[1];
          w[2] = q[2];
        }
      }
      // any lane's best bounds every lane's search: a triangle farther
      // than it (strictly) cannot be the first minimum either
      double m = bd;
      for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off));
      pb = m;
    }
  }
  if (!live) return;
  // smallest distance over the wave, ties to the lowest triangle index
  for (int off = 32; off > 0; off >>= 1) {
    const double od = __shfl_xor(bd, off);
    const int64_t ot = __shfl_xor(bt, off);
    const double o0 = __shfl_xor(w[0], off), o1 = __shfl_xor(w[1], off), o2 = __shfl_xor(w[2], off);
    if (od < bd || (od == bd && ot < bt)) {
      bd = od;
      bt = ot;
      w[0] = o0;
      w[1] = o1;
      w[2] = o2;
    }
  }
  bool valid = false;
  double qv = -1.0;
  if (lane < kQ - 1) {
    const int ip = lane + 1;
    const int o = Opp(ip);
    const int ex = Cx(o), ey = Cy(o), ez = Cz(o);
    const int ni = i + ex, nj = j + ey, nk = k + ez;
    const uint8_t* mat = material + static_cast<size_t>(b) * n * n * n;
    if (ni >= 0 && ni < n && nj >= 0 && nj < n && nk >= 0 && nk < n && SolidLike(Mat(mat, n, ni, nj, nk))) {
      const double dir_x = static_cast<double>(ex) * dx;
      const double dir_y = static_cast<double>(ey) * dx;
      const double dir_z = static_cast<double>(ez) * dx;
      const double dn2 = dir_x * dir_x + dir_y * dir_y + dir_z * dir_z;
      if (dn2 > 1.0e-30) {
        const double tx = w[0] - p[0], ty = w[1] - p[1], tz = w[2] - p[2];
        const double proj = (tx * dir_x + ty * dir_y + tz * dir_z) / dn2;
        if (proj > 1.0e-8 && proj < 1.0 - 1.0e-8) {
          valid = true;
          qv = proj;
        }
      }
    }
    qd[ci * (kQ - 1) + lane] = qv;
  }
  const uint64_t m = __ballot(valid);
  if (lane == 0) cnt[ci] = __popcll(m);
}

// One wave per candidate: its links at offset[cand], population order.
__global__ __launch_bounds__(kLinkThreads) void KEmit(const int2* cand, int64_t ncand, const double* qd,
                                                      const int64_t* offset, int32_t* links, double* q) {
  const int lane = threadIdx.x & 63;
  const int64_t ci = static_cast<int64_t>(blockIdx.x) * (kLinkThreads / 64) + (threadIdx.x >> 6);
  if (ci >= ncand) return;
  const double v = lane < kQ - 1 ? qd[ci * (kQ - 1) + lane] : -1.0;
  const bool valid = v >= 0.0;
  const uint64_t m = __ballot(valid);
  if (!valid) return;
  const int64_t at = offset[ci] + __popcll(m & ((uint64_t(1) << lane) - 1));
  const int2 e = cand[ci];
  links[3 * at] = e.x;
  links[3 * at + 1] = e.y;
  links[3 * at + 2] = lane + 1;
  q[at] = v;
}

struct DevBuf {
  std::vector<void*> ptrs;
  template <typename P>
  P* Get(size_t count) {
    void* p = nullptr;
    LINK_CHECK(hipMalloc(&p, (count == 0 ? 1 : count) * sizeof(P)));
    ptrs.push_back(p);
    return static_cast<P*>(p);
  }
  // Teardown: a failed release cannot be reported from here.
  ~DevBuf() {
    for (void* p : ptrs) (void)hipFree(p);
  }
};

struct Stream {
  hipStream_t s = nullptr;
  Stream() { LINK_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);  // teardown
  }
};

void BuildLinks(int32_t nb, int32_t n, const double* bounds, const uint8_t* material, int64_t ntri,
                const double* tris, int64_t capacity, int64_t* num_links, int32_t* links, double* q) {
  *num_links = 0;
  if (ntri == 0 || nb == 0) return;  // soup.empty(): no links (bouzidi_link_data.cpp:147-149)
  const size_t ncell = static_cast<size_t>(n) * n * n;
  Stream st;
  DevBuf mem;
  double* d_bounds = mem.Get<double>(6 * static_cast<size_t>(nb));
  uint8_t* d_mat = mem.Get<uint8_t>(ncell * nb);
  double* d_tris = mem.Get<double>(9 * static_cast<size_t>(ntri));
  int* d_counts = mem.Get<int>(nb);
  int64_t* d_boff = mem.Get<int64_t>(static_cast<size_t>(nb) + 1);
  LINK_CHECK(hipMemcpyAsync(d_bounds, bounds, 6 * sizeof(double) * nb, hipMemcpyHostToDevice, st.s));
  LINK_CHECK(hipMemcpyAsync(d_mat, material, ncell * nb, hipMemcpyHostToDevice, st.s));
  LINK_CHECK(hipMemcpyAsync(d_tris, tris, 9 * sizeof(double) * ntri, hipMemcpyHostToDevice, st.s));
  KCandidates<false><<<nb, kLinkThreads, 0, st.s>>>(n, d_mat, d_counts, nullptr, nullptr);
  LINK_CHECK(hipGetLastError());
  KScan<<<1, 1024, 0, st.s>>>(d_counts, nb, d_boff);
  LINK_CHECK(hipGetLastError());
  int64_t ncand = 0;
  LINK_CHECK(hipMemcpyAsync(&ncand, d_boff + nb, sizeof(int64_t), hipMemcpyDeviceToHost, st.s));
  LINK_CHECK(hipStreamSynchronize(st.s));
  if (ncand == 0) return;
  int2* d_cand = mem.Get<int2>(ncand);
  KCandidates<true><<<nb, kLinkThreads, 0, st.s>>>(n, d_mat, nullptr, d_boff, d_cand);
  LINK_CHECK(hipGetLastError());
  double* d_qd = mem.Get<double>(static_cast<size_t>(ncand) * (kQ - 1));
  int* d_cnt = mem.Get<int>(ncand);
  int64_t* d_loff = mem.Get<int64_t>(static_cast<size_t>(ncand) + 1);
  const int64_t waves_per_wg = kLinkThreads / 64;
  const unsigned grid = static_cast<unsigned>((ncand + waves_per_wg - 1) / waves_per_wg);
  KNearest<<<grid, kLinkThreads, 0, st.s>>>(d_cand, ncand, n, d_bounds, d_mat, d_tris, ntri, d_qd, d_cnt);
  LINK_CHECK(hipGetLastError());
  KScan<<<1, 1024, 0, st.s>>>(d_cnt, ncand, d_loff);
  LINK_CHECK(hipGetLastError());
  int64_t total = 0;
  LINK_CHECK(hipMemcpyAsync(&total, d_loff + ncand, sizeof(int64_t), hipMemcpyDeviceToHost, st.s));
  LINK_CHECK(hipStreamSynchronize(st.s));
  *num_links = total;
  if (links == nullptr && q == nullptr) return;  // count query
  if (capacity < total) throw std::invalid_argument("link capacity smaller than the number of links");
  if (total == 0) return;
  int32_t* d_links = mem.Get<int32_t>(3 * static_cast<size_t>(total));
  double* d_q = mem.Get<double>(total);
  KEmit<<<grid, kLinkThreads, 0, st.s>>>(d_cand, ncand, d_qd, d_loff, d_links, d_q);
  LINK_CHECK(hipGetLastError());
  LINK_CHECK(hipMemcpyAsync(links, d_links, 3 * sizeof(int32_t) * total, hipMemcpyDeviceToHost, st.s));
  LINK_CHECK(hipMemcpyAsync(q, d_q, sizeof(double) * total, hipMemcpyDeviceToHost, st.s));
  LINK_CHECK(hipStreamSynchronize(st.s));
}

// ---- no-allocation form: every scratch buffer of the build lives in one
// file-scope device array per GPU, sized for at most kNoAllocMaxBlocks
// blocks and kNoAllocMaxCandidates candidate cells.
constexpr int64_t kNoAllocMaxBlocks = 16384;
constexpr int64_t kNoAllocMaxCandidates = 65536;

constexpr size_t Align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }
constexpr size_t kScrCounts = 0;
constexpr size_t kScrBoff = kScrCounts + Align256(sizeof(int) * kNoAllocMaxBlocks);
constexpr size_t kScrCand = kScrBoff + Align256(sizeof(int64_t) * (kNoAllocMaxBlocks + 1));
constexpr size_t kScrQd = kScrCand + Align256(sizeof(int2) * kNoAllocMaxCandidates);
constexpr size_t kScrCnt = kScrQd + Align256(sizeof(double) * (kQ - 1) * kNoAllocMaxCandidates);
constexpr size_t kScrLoff = kScrCnt + Align256(sizeof(int) * kNoAllocMaxCandidates);
constexpr size_t kScratchBytes = kScrLoff + Align256(sizeof(int64_t) * (kNoAllocMaxCandidates + 1));

}  // namespace

// ~10.9 MB per GPU, held until the process ends.  In host code this name is
// only a symbol handle: its device address comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current GPU, looked up once per device.
unsigned char* ScratchOnCurrentDevice() {
  int dev = 0;
  LINK_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    LINK_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[dev];
}

void BuildLinksNoAlloc(int32_t nb, int32_t n, const double* d_bounds, const uint8_t* d_mat, int64_t ntri,
                       const double* d_tris, int64_t capacity, int64_t* num_links, int32_t* d_links, double* d_q,
                       hipStream_t s) {
  *num_links = 0;
  if (nb > kNoAllocMaxBlocks) throw std::invalid_argument("more blocks than the no-allocation scratch holds (16384)");
  if (ntri == 0 || nb == 0) return;  // soup.empty(): no links (bouzidi_link_data.cpp:147-149)
  unsigned char* base = ScratchOnCurrentDevice();
  int* d_counts = reinterpret_cast<int*>(base + kScrCounts);
  int64_t* d_boff = reinterpret_cast<int64_t*>(base + kScrBoff);
  int2* d_cand = reinterpret_cast<int2*>(base + kScrCand);
  double* d_qd = reinterpret_cast<double*>(base + kScrQd);
  int* d_cnt = reinterpret_cast<int*>(base + kScrCnt);
  int64_t* d_loff = reinterpret_cast<int64_t*>(base + kScrLoff);
  // every scratch word read below is written first in this call
  KCandidates<false><<<nb, kLinkThreads, 0, s>>>(n, d_mat, d_counts, nullptr, nullptr);
  LINK_CHECK(hipGetLastError());
  KScan<<<1, 1024, 0, s>>>(d_counts, nb, d_boff);
  LINK_CHECK(hipGetLastError());
  int64_t ncand = 0;
  LINK_CHECK(hipMemcpyAsync(&ncand, d_boff + nb, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  LINK_CHECK(hipStreamSynchronize(s));
  if (ncand == 0) return;
  if (ncand > kNoAllocMaxCandidates)
    throw std::invalid_argument("more candidate cells than the no-allocation scratch holds (65536)");
  KCandidates<true><<<nb, kLinkThreads, 0, s>>>(n, d_mat, nullptr, d_boff, d_cand);
  LINK_CHECK(hipGetLastError());
  const int64_t waves_per_wg = kLinkThreads / 64;
  const unsigned grid = static_cast<unsigned>((ncand + waves_per_wg - 1) / waves_per_wg);
  KNearest<<<grid, kLinkThreads, 0, s>>>(d_cand, ncand, n, d_bounds, d_mat, d_tris, ntri, d_qd, d_cnt);
  LINK_CHECK(hipGetLastError());
  KScan<<<1, 1024, 0, s>>>(d_cnt, ncand, d_loff);
  LINK_CHECK(hipGetLastError());
  int64_t total = 0;
  LINK_CHECK(hipMemcpyAsync(&total, d_loff + ncand, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  LINK_CHECK(hipStreamSynchronize(s));
  *num_links = total;
  if (d_links == nullptr && d_q == nullptr) return;  // count query
  if (capacity < total) throw std::invalid_argument("link capacity smaller than the number of links");
  if (total == 0) return;
  KEmit<<<grid, kLinkThreads, 0, s>>>(d_cand, ncand, d_qd, d_loff, d_links, d_q);
  LINK_CHECK(hipGetLastError());
  LINK_CHECK(hipStreamSynchronize(s));
}

}  // namespace
}  // namespace octlb

extern "C" int octlb_build_links_noalloc(int32_t num_blocks, int32_t n, const double* bounds,
                                         const uint8_t* material, int64_t num_triangles, const double* triangles,
                                         int64_t capacity, int64_t* num_links, int32_t* links, double* q,
                                         void* stream) {
  using namespace octlb;
  if (!num_links) return SetError(OCTLB_EINVAL, "null num_links");
  *num_links = 0;
  if (num_blocks < 0 || n < 1 || n > 1024 || num_triangles < 0 || capacity < 0)
    return SetError(OCTLB_EINVAL, "bad link-build sizes");
  if ((num_blocks > 0 && (!bounds || !material)) || (num_triangles > 0 && !triangles))
    return SetError(OCTLB_EINVAL, "null link-build input");
  if ((links == nullptr) != (q == nullptr)) return SetError(OCTLB_EINVAL, "links and q must both be given");
  try {
    BuildLinksNoAlloc(num_blocks, n, bounds, material, num_triangles, triangles, capacity, num_links, links, q,
                      static_cast<hipStream_t>(stream));
  } catch (const std::invalid_argument& e) {
    return SetError(OCTLB_EINVAL, e.what());
  } catch (const std::exception& e) {
    return SetError(OCTLB_EHIP, e.what());
  }
  return OCTLB_OK;
}

extern "C" int octlb_build_links(int32_t num_blocks, int32_t n, const double* bounds, const uint8_t* material,
                                 int64_t num_triangles, const double* triangles, int64_t capacity,
                                 int64_t* num_links, int32_t* links, double* q) {
  using namespace octlb;
  if (!num_links) return SetError(OCTLB_EINVAL, "null num_links");
  *num_links = 0;
  if (num_blocks < 0 || n < 1 || n > 1024 || num_triangles < 0 || capacity < 0)
    return SetError(OCTLB_EINVAL, "bad link-build sizes");
  if ((num_blocks > 0 && (!bounds || !material)) || (num_triangles > 0 && !triangles))
    return SetError(OCTLB_EINVAL, "null link-build input");
  if ((links == nullptr) != (q == nullptr)) return SetError(OCTLB_EINVAL, "links and q must both be given");
  try {
    BuildLinks(num_blocks, n, bounds, material, num_triangles, triangles, capacity, num_links, links, q);
  } catch (const std::invalid_argument& e) {
    return SetError(OCTLB_EINVAL, e.what());
  } catch (const std::bad_alloc&) {
    return SetError(OCTLB_ENOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return SetError(OCTLB_EHIP, e.what());
  }
  return OCTLB_OK;
}
