// Host-side brick forest for the AMR block topology.
//
// The reference keeps its octree in p4est (src/mesh/forest/octree_forest.cpp)
// and enumerates block adjacency with one p8est_iterate pass
// (src/mesh/topology/face_pair_list.cpp:176-337, 480-558).  p4est is not part
// of this build, so this file provides the same public contract for brick
// domains:
//   * leaves carry (tree, level, integer lower corner) in a fixed-depth
//     integer space, and are ordered tree by tree (trees in Morton order of
//     their brick coordinates), Morton order inside a tree
//     (child id = x + 2y + 4z, as in p8est);
//   * refine(bounds predicate, max_level) is recursive, like p8est_refine_ext
//     with recursive=1 (octree_forest.cpp:248-261);
//   * balance() enforces 2:1 across faces only (P8EST_CONNECT_FACE,
//     octree_forest.cpp:263-271);
//   * the face lists are SameLevelFace / CoarseFineFace / TreeBoundaryFace
//     (face_pair_list.h:13-45).  One rank owns every octant, so there are no
//     cross-rank edges and all remote ranks are 0.
// Enumeration order is octant order then face 0..5; coarse-fine faces are
// stably sorted by coarse level so that each coarse level occupies one
// contiguous range of the coupling plan (LevelCoupler::LevelRangeBegin/End,
// level_coupler.cpp:309-323, assume that).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "octlb_gpu.h"
#include "octlb_internal.h"

namespace {

constexpr int kMaxLevel = 12;  // integer depth of one tree

struct Leaf {
  int32_t tree;
  int32_t level;
  int64_t x, y, z;  // lower corner in units of 2^-kMaxLevel tree edges (global)
};

inline int64_t LeafSize(int level) { return int64_t{1} << (kMaxLevel - level); }

inline uint64_t Key(int level, int64_t x, int64_t y, int64_t z) {
  return (static_cast<uint64_t>(level) << 57) | (static_cast<uint64_t>(x) << 38) |
         (static_cast<uint64_t>(y) << 19) | static_cast<uint64_t>(z);
}

inline uint64_t Spread3(uint64_t v) {
  uint64_t r = 0;
  for (int b = 0; b < 21; ++b) r |= ((v >> b) & 1u) << (3 * b);
  return r;
}

inline uint64_t Morton(uint64_t x, uint64_t y, uint64_t z) {
  return Spread3(x) | (Spread3(y) << 1) | (Spread3(z) << 2);
}

}  // namespace

struct octlb_forest {
  double domain[6];
  int32_t bricks[3];
  std::vector<int32_t> tree_rank;  // brick-linear tree index -> Morton rank
  std::vector<Leaf> leaves;        // sorted
  std::unordered_map<uint64_t, int32_t> index;

  // topology storage
  std::vector<int32_t> level;
  std::vector<double> bounds;
  std::vector<int32_t> same_level;
  std::vector<octlb_cf_face> coarse_fine;
  std::vector<int32_t> tree_boundary;

  int64_t extent(int axis) const { return int64_t{bricks[axis]} << kMaxLevel; }

  int32_t TreeOf(int64_t x, int64_t y, int64_t z) const {
    const int64_t tx = x >> kMaxLevel, ty = y >> kMaxLevel, tz = z >> kMaxLevel;
    return static_cast<int32_t>((tz * bricks[1] + ty) * bricks[0] + tx);
  }

  void Sort() {
    std::sort(leaves.begin(), leaves.end(), [&](const Leaf& a, const Leaf& b) {
      const int32_t ra = tree_rank[a.tree], rb = tree_rank[b.tree];
      if (ra != rb) return ra < rb;
      const int64_t m = (int64_t{1} << kMaxLevel) - 1;
      return Morton(a.x & m, a.y & m, a.z & m) < Morton(b.x & m, b.y & m, b.z & m);
    });
    index.clear();
    for (size_t i = 0; i < leaves.size(); ++i) {
      const Leaf& l = leaves[i];
      index[Key(l.level, l.x, l.y, l.z)] = static_cast<int32_t>(i);
    }
  }

  bool Inside(int64_t x, int64_t y, int64_t z) const {
    return x >= 0 && y >= 0 && z >= 0 && x < extent(0) && y < extent(1) && z < extent(2);
  }

  // Leaf containing integer point p, or -1 outside the domain.
  int32_t Locate(int64_t x, int64_t y, int64_t z) const {
    if (!Inside(x, y, z)) return -1;
    for (int l = 0; l <= kMaxLevel; ++l) {
      const int64_t s = LeafSize(l);
      auto it = index.find(Key(l, (x / s) * s, (y / s) * s, (z / s) * s));
      if (it != index.end()) return it->second;
    }
    return -1;
  }

  void Bounds(const Leaf& l, double out[6]) const {
    const int64_t s = LeafSize(l.level);
    const int64_t lo[3] = {l.x, l.y, l.z};
    for (int a = 0; a < 3; ++a) {
      const double ext = domain[3 + a] - domain[a];
      const double denom = static_cast<double>(extent(a));
      out[a] = domain[a] + ext * (static_cast<double>(lo[a]) / denom);
      out[3 + a] = domain[a] + ext * (static_cast<double>(lo[a] + s) / denom);
    }
  }

  void RefineLeaf(const Leaf& l, std::vector<Leaf>* out) const {
    const int64_t h = LeafSize(l.level + 1);
    for (int c = 0; c < 8; ++c) {
      Leaf k = l;
      k.level = l.level + 1;
      k.x = l.x + ((c & 1) ? h : 0);
      k.y = l.y + ((c & 2) ? h : 0);
      k.z = l.z + ((c & 4) ? h : 0);
      out->push_back(k);
    }
  }
};

namespace {

// Point just outside face `dir` of leaf l (at its lower tangential corner).
void FaceProbe(const Leaf& l, int dir, int64_t* p) {
  const int64_t s = LeafSize(l.level);
  p[0] = l.x;
  p[1] = l.y;
  p[2] = l.z;
  const int axis = dir / 2;
  p[axis] += (dir & 1) ? s : -1;
}

}  // namespace

extern "C" {

int octlb_forest_create(const double domain[6], int32_t bx, int32_t by, int32_t bz,
                        octlb_forest** out) {
  try {
    if (domain == nullptr || out == nullptr) return octlb::SetError(OCTLB_EINVAL, "null argument");
    if (bx < 1 || by < 1 || bz < 1 || bx > 64 || by > 64 || bz > 64)
      return octlb::SetError(OCTLB_EINVAL, "brick dimensions must be in [1,64]");
    auto* f = new octlb_forest();
    std::memcpy(f->domain, domain, sizeof(f->domain));
    f->bricks[0] = bx;
    f->bricks[1] = by;
    f->bricks[2] = bz;
    const int nt = bx * by * bz;
    std::vector<std::pair<uint64_t, int32_t>> order;
    for (int tz = 0; tz < bz; ++tz)
      for (int ty = 0; ty < by; ++ty)
        for (int tx = 0; tx < bx; ++tx)
          order.push_back({Morton(tx, ty, tz), (tz * by + ty) * bx + tx});
    std::sort(order.begin(), order.end());
    f->tree_rank.assign(nt, 0);
    for (int r = 0; r < nt; ++r) f->tree_rank[order[r].second] = r;
    for (int tz = 0; tz < bz; ++tz)
      for (int ty = 0; ty < by; ++ty)
        for (int tx = 0; tx < bx; ++tx) {
          Leaf l{(tz * by + ty) * bx + tx, 0, int64_t{tx} << kMaxLevel,
                 int64_t{ty} << kMaxLevel, int64_t{tz} << kMaxLevel};
          f->leaves.push_back(l);
        }
    f->Sort();
    *out = f;
    return OCTLB_OK;
  } catch (const std::exception& e) {
    return octlb::SetError(OCTLB_ENOMEM, e.what());
  }
}

void octlb_forest_destroy(octlb_forest* forest) { delete forest; }

int octlb_forest_refine_box(octlb_forest* f, const double box[6], int32_t max_level) {
  if (f == nullptr || box == nullptr) return octlb::SetError(OCTLB_EINVAL, "null argument");
  if (max_level < 0 || max_level > kMaxLevel - 1)
    return octlb::SetError(OCTLB_EINVAL, "max_level out of range");
  std::vector<Leaf> work = f->leaves, done;
  while (!work.empty()) {
    std::vector<Leaf> next;
    for (const Leaf& l : work) {
      double b[6];
      f->Bounds(l, b);
      const bool overlap = !(b[3] < box[0] || b[0] > box[3] || b[4] < box[1] ||
                             b[1] > box[4] || b[5] < box[2] || b[2] > box[5]);
      if (l.level < max_level && overlap) {
        f->RefineLeaf(l, &next);
      } else {
        done.push_back(l);
      }
    }
    work.swap(next);
  }
  f->leaves.swap(done);
  f->Sort();
  return OCTLB_OK;
}

int octlb_forest_refine_marked(octlb_forest* f, const uint8_t* mark, int64_t count) {
  if (f == nullptr || mark == nullptr) return octlb::SetError(OCTLB_EINVAL, "null argument");
  if (count != static_cast<int64_t>(f->leaves.size()))
    return octlb::SetError(OCTLB_EINVAL, "mark count does not match octant count");
  std::vector<Leaf> out;
  for (size_t i = 0; i < f->leaves.size(); ++i) {
    const Leaf& l = f->leaves[i];
    if (mark[i] && l.level < kMaxLevel - 1) {
      f->RefineLeaf(l, &out);
    } else {
      out.push_back(l);
    }
  }
  f->leaves.swap(out);
  f->Sort();
  return OCTLB_OK;
}

int octlb_forest_balance(octlb_forest* f) {
  if (f == nullptr) return octlb::SetError(OCTLB_EINVAL, "null forest");
  for (;;) {
    std::vector<uint8_t> mark(f->leaves.size(), 0);
    bool any = false;
    for (const Leaf& l : f->leaves) {
      for (int dir = 0; dir < 6; ++dir) {
        int64_t p[3];
        FaceProbe(l, dir, p);
        const int32_t n = f->Locate(p[0], p[1], p[2]);
        if (n < 0) continue;
        if (f->leaves[n].level < l.level - 1) {
          mark[n] = 1;
          any = true;
        }
      }
    }
    if (!any) break;
    std::vector<Leaf> out;
    for (size_t i = 0; i < f->leaves.size(); ++i) {
      if (mark[i]) {
        f->RefineLeaf(f->leaves[i], &out);
      } else {
        out.push_back(f->leaves[i]);
      }
    }
    f->leaves.swap(out);
    f->Sort();
  }
  return OCTLB_OK;
}

int32_t octlb_forest_num_octants(const octlb_forest* f) {
  return f == nullptr ? 0 : static_cast<int32_t>(f->leaves.size());
}

int octlb_forest_topology(octlb_forest* f, int32_t n, int32_t halo, octlb_topology* out) {
  if (f == nullptr || out == nullptr) return octlb::SetError(OCTLB_EINVAL, "null argument");
  if (n < 2 || (n % 2) != 0) return octlb::SetError(OCTLB_EINVAL, "block edge n must be even");
  if (halo < 0) return octlb::SetError(OCTLB_EINVAL, "halo must be >= 0");
  const int32_t nb = static_cast<int32_t>(f->leaves.size());
  f->level.assign(nb, 0);
  f->bounds.assign(static_cast<size_t>(nb) * 6, 0.0);
  f->same_level.clear();
  f->coarse_fine.clear();
  f->tree_boundary.clear();
  for (int32_t i = 0; i < nb; ++i) {
    f->level[i] = f->leaves[i].level;
    f->Bounds(f->leaves[i], &f->bounds[static_cast<size_t>(i) * 6]);
  }
  for (int32_t i = 0; i < nb; ++i) {
    const Leaf& l = f->leaves[i];
    const int64_t s = LeafSize(l.level);
    for (int dir = 0; dir < 6; ++dir) {
      int64_t p[3];
      FaceProbe(l, dir, p);
      const int32_t nbr = f->Locate(p[0], p[1], p[2]);
      if (nbr < 0) {
        f->tree_boundary.push_back(i);
        f->tree_boundary.push_back(dir);
        continue;
      }
      const Leaf& m = f->leaves[nbr];
      if (m.level == l.level) {
        f->same_level.push_back(i);
        f->same_level.push_back(dir);
        f->same_level.push_back(nbr);
      } else if (m.level == l.level + 1) {
        const int axis = dir / 2;
        const int t0 = (axis + 1) % 3, t1 = (axis + 2) % 3;
        const int64_t h = s / 2;
        std::vector<int32_t> fines;
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b) {
            int64_t q[3] = {p[0], p[1], p[2]};
            q[t0] += a * h;
            q[t1] += b * h;
            const int32_t fi = f->Locate(q[0], q[1], q[2]);
            if (fi < 0 || f->leaves[fi].level != l.level + 1)
              return octlb::SetError(OCTLB_EINVAL, "forest is not 2:1 face balanced");
            fines.push_back(fi);
          }
        std::sort(fines.begin(), fines.end());
        octlb_cf_face cf{};
        cf.coarse_id = i;
        for (int k = 0; k < 4; ++k) cf.fine_ids[k] = fines[k];
        cf.normal = dir;
        cf.coarse_level = l.level;
        cf.fine_level = l.level + 1;
        std::memcpy(cf.coarse_bounds, &f->bounds[static_cast<size_t>(i) * 6], sizeof(double) * 6);
        for (int k = 0; k < 4; ++k)
          std::memcpy(cf.fine_bounds[k], &f->bounds[static_cast<size_t>(fines[k]) * 6],
                      sizeof(double) * 6);
        f->coarse_fine.push_back(cf);
      } else if (m.level > l.level + 1) {
        return octlb::SetError(OCTLB_EINVAL, "forest is not 2:1 face balanced");
      }
      // m.level == l.level - 1: recorded from the coarse side.
    }
  }
  std::stable_sort(f->coarse_fine.begin(), f->coarse_fine.end(),
                   [](const octlb_cf_face& a, const octlb_cf_face& b) {
                     return a.coarse_level < b.coarse_level;
                   });
  out->num_blocks = nb;
  out->n = n;
  out->halo = halo;
  out->level = f->level.data();
  out->bounds = f->bounds.data();
  out->num_same_level = static_cast<int32_t>(f->same_level.size() / 3);
  out->same_level = f->same_level.data();
  out->num_coarse_fine = static_cast<int32_t>(f->coarse_fine.size());
  out->coarse_fine = f->coarse_fine.data();
  out->num_tree_boundary = static_cast<int32_t>(f->tree_boundary.size() / 2);
  out->tree_boundary = f->tree_boundary.data();
  return OCTLB_OK;
}

}  // extern "C"
