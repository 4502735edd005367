/*
 * Source-compatible device TimeLoop for the reference's C++ solver objects.
 *
 * A maintainer switches a case to the MI355X path by replacing
 *   octlb::TimeLoop                          -> octlb::gpu::TimeLoop
 *   octlb::ConcreteDomainBoundaryHandler     -> octlb::gpu::ConcreteDomainBoundaryHandler
 *   boundary::UniformInletProfile (and the Poiseuille / channel profiles)
 *                                            -> octlb::gpu::UniformInletProfile ...
 * with unchanged constructor arguments (time_loop.h:49-54,
 * domain_boundary_handler.h:105-111, inlet_velocity_field.h:47-56, 103-104,
 * 136-147).  The replacement handler and inlet classes derive from the
 * reference's own and only remember their constructor arguments, which the
 * reference keeps private; the device path needs them to flatten the
 * boundary specs once.
 *
 * gpu::TimeLoop flattens, once at construction ("once per adapt"):
 *   - levels and bounds from OctreeForest::quadrant_level / quadrant_bounds;
 *   - same-level faces from the block geometry (on one rank these are the
 *     FacePairList::same_level_faces pairs: same level, touching faces);
 *   - coarse-fine faces from LevelCoupler::coupling_plan(), in plan order;
 *   - tree-boundary faces from the handler (or from the geometry);
 *   - BcKind bytes, Bouzidi q (BouzidiLinkData::q_frac != 1) and the
 *     halo-inclusive populations from the BlockCollection;
 * and then owns one device handle (liboctlb_gpu.so).  advance_one() steps
 * the device state; the host lattices are stale until sync_to_host().
 *
 * Errors keep the reference's convention: std::runtime_error with the
 * library's message (octlb_last_error), std::invalid_argument for objects
 * the device path cannot take.
 *
 * Needs the reference's include root on the include path (-I<reference>,
 * -I<reference>/src/solver/lbm for the vendored OpenLB headers) and this
 * repository's include/ directory.
 */
#ifndef OCTLB_GPU_TIME_LOOP_H_
#define OCTLB_GPU_TIME_LOOP_H_

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "octlb_gpu.h"
#include "src/solver/lbm/boundary/inlet_velocity_field.h"
#include "src/solver/lbm/bouzidi_link_data.h"
#include "src/solver/lbm/domain_boundary_handler.h"
#include "src/solver/lbm/time_loop/time_loop.h"

namespace octlb {
namespace gpu {

inline void Check(int rc) {
  if (rc != OCTLB_OK) throw std::runtime_error(octlb_last_error());
}

// ------------------------------------------------------------ inlet fields
/** What the device needs of an inlet profile: its C-ABI description. */
class DeviceInletField {
 public:
  virtual ~DeviceInletField() = default;
  /** Fills the inlet part of a spec (inlet type, face, direction, peak,
   *  extents, ramp). */
  virtual void describe(octlb_bc_spec* s) const = 0;
};

/** boundary::UniformInletProfile (inlet_velocity_field.h:131-176). */
class UniformInletProfile : public boundary::UniformInletProfile, public DeviceInletField {
 public:
  UniformInletProfile(std::array<double, 3> u_const, double ramp_end_t = 0.0)
      : boundary::UniformInletProfile(u_const, ramp_end_t), u_const_(u_const), ramp_end_t_(ramp_end_t) {}
  void describe(octlb_bc_spec* s) const override {
    s->inlet = OCTLB_INLET_UNIFORM;
    for (int d = 0; d < 3; ++d) s->inlet_u[d] = u_const_[d];
    s->ramp_end_t = ramp_end_t_;
  }

 private:
  std::array<double, 3> u_const_;
  double ramp_end_t_;
};

/** boundary::PoiseuilleInletProfile (inlet_velocity_field.h:47-129). */
class PoiseuilleInletProfile : public boundary::PoiseuilleInletProfile, public DeviceInletField {
 public:
  PoiseuilleInletProfile(FaceDir face_dir, int n_inplane0, int n_inplane1, double u_peak,
                         std::array<double, 3> flow_dir, double ramp_end_t = 0.0)
      : boundary::PoiseuilleInletProfile(face_dir, n_inplane0, n_inplane1, u_peak, flow_dir, ramp_end_t),
        face_(face_dir), n0_(n_inplane0), n1_(n_inplane1), u_peak_(u_peak), dir_(flow_dir),
        ramp_end_t_(ramp_end_t) {}
  void describe(octlb_bc_spec* s) const override {
    s->inlet = OCTLB_INLET_POISEUILLE;
    s->inlet_face = static_cast<int32_t>(face_);
    for (int d = 0; d < 3; ++d) s->inlet_u[d] = dir_[d];
    s->u_peak = u_peak_;
    s->n_inplane0 = n0_;
    s->n_inplane1 = n1_;
    s->ramp_end_t = ramp_end_t_;
  }

 private:
  FaceDir face_;
  int n0_, n1_;
  double u_peak_;
  std::array<double, 3> dir_;
  double ramp_end_t_;
};

/** boundary::ChannelPoiseuilleInletProfile (inlet_velocity_field.h:178-270). */
class ChannelPoiseuilleInletProfile : public boundary::ChannelPoiseuilleInletProfile, public DeviceInletField {
 public:
  ChannelPoiseuilleInletProfile(FaceDir face_dir, double in0_lo, double in0_hi, double in1_lo, double in1_hi,
                                double u_peak, std::array<double, 3> flow_dir, double ramp_end_t = 0.0)
      : boundary::ChannelPoiseuilleInletProfile(face_dir, in0_lo, in0_hi, in1_lo, in1_hi, u_peak, flow_dir,
                                                ramp_end_t),
        face_(face_dir), lo0_(in0_lo), hi0_(in0_hi), lo1_(in1_lo), hi1_(in1_hi), u_peak_(u_peak), dir_(flow_dir),
        ramp_end_t_(ramp_end_t) {}
  void describe(octlb_bc_spec* s) const override {
    s->inlet = OCTLB_INLET_CHANNEL;
    s->inlet_face = static_cast<int32_t>(face_);
    for (int d = 0; d < 3; ++d) s->inlet_u[d] = dir_[d];
    s->u_peak = u_peak_;
    s->lo0 = lo0_;
    s->hi0 = hi0_;
    s->lo1 = lo1_;
    s->hi1 = hi1_;
    s->ramp_end_t = ramp_end_t_;
  }

 private:
  FaceDir face_;
  double lo0_, hi0_, lo1_, hi1_, u_peak_;
  std::array<double, 3> dir_;
  double ramp_end_t_;
};

// --------------------------------------------------------- boundary handler
/** ConcreteDomainBoundaryHandler (domain_boundary_handler.h:103-133) that
 *  also keeps its tree-boundary faces and specs for the device path.  On
 *  the host it behaves exactly as the reference's handler. */
class ConcreteDomainBoundaryHandler : public octlb::ConcreteDomainBoundaryHandler {
 public:
  ConcreteDomainBoundaryHandler(BlockCollection<DomainBoundaryLattice>& blocks,
                                const std::vector<TreeBoundaryFace>& faces, const std::vector<DomainBcSpec>& specs,
                                int nx, int ny, int nz, double omega = 1.0,
                                OverlapPaddingMode padding_mode = OverlapPaddingMode::kHybrid)
      : octlb::ConcreteDomainBoundaryHandler(blocks, faces, specs, nx, ny, nz, omega, padding_mode),
        faces_(faces), specs_(specs), padding_mode_(padding_mode) {}

  const std::vector<TreeBoundaryFace>& tree_boundary_faces() const { return faces_; }
  const std::vector<DomainBcSpec>& specs() const { return specs_; }
  OverlapPaddingMode padding_mode() const { return padding_mode_; }

 private:
  std::vector<TreeBoundaryFace> faces_;
  std::vector<DomainBcSpec> specs_;
  OverlapPaddingMode padding_mode_;
};

// ---------------------------------------------------------------- TimeLoop
/** Device replacement of octlb::TimeLoop (time_loop.h:49-99). */
class TimeLoop {
 public:
  /** Same arguments as octlb::TimeLoop; `precision` (64 or 32: FP32
   *  shifted populations) is the one addition, defaulted to the reference's
   *  FP64. */
  TimeLoop(const OctreeForest& forest, BlockCollection<TimeLoopLattice>& blocks,
           GhostSchedule<TimeLoopLattice>& ghosts, LevelCoupler& coupler, DomainBoundaryHandler& domain_bc,
           double omega, bool use_const_rho_bgk = false,
           ConstRhoStatsScope const_rho_stats_scope = ConstRhoStatsScope::kFluidAndBoundary, int precision = 64)
      : blocks_(blocks) {
    (void)ghosts;  // its faces are the geometric same-level pairs rebuilt below
    nb_ = static_cast<int>(forest.local_num_octants());
    if (nb_ < 1 || blocks.size() != nb_) throw std::invalid_argument("gpu::TimeLoop: forest and blocks disagree");
    n_ = blocks[0].nx();
    halo_ = blocks[0].halo_width();
    for (int b = 0; b < nb_; ++b)
      if (blocks[b].nx() != n_ || blocks[b].ny() != n_ || blocks[b].nz() != n_ || blocks[b].halo_width() != halo_)
        throw std::invalid_argument("gpu::TimeLoop: blocks must be cubes of one size and halo");
    FlattenForest(forest);
    FlattenCoupling(forest, coupler);
    const octlb_params p = FlattenHandler(domain_bc, omega, use_const_rho_bgk, const_rho_stats_scope, precision);
    const octlb_topology topo{nb_,
                              n_,
                              halo_,
                              level_.data(),
                              bounds_.data(),
                              static_cast<int32_t>(same_.size() / 3),
                              same_.data(),
                              static_cast<int32_t>(cf_.size()),
                              cf_.data(),
                              static_cast<int32_t>(tree_.size() / 2),
                              tree_.data()};
    Check(octlb_dev_create(&topo, &p, &dev_));
    try {
      UploadBlocks();
    } catch (...) {
      octlb_dev_destroy(dev_);
      throw;
    }
    max_level_ = *std::max_element(level_.begin(), level_.end());
    counters_.collide.assign(static_cast<std::size_t>(max_level_ + 1), 0);
    counters_.stream.assign(static_cast<std::size_t>(max_level_ + 1), 0);
  }
  ~TimeLoop() { octlb_dev_destroy(dev_); }
  TimeLoop(const TimeLoop&) = delete;
  TimeLoop& operator=(const TimeLoop&) = delete;

  /** TimeLoop::advance_one (time_loop.h:242-246) on the device. */
  void advance_one() {
    Check(octlb_dev_advance(dev_, 1));
    Count(0);
  }
  /** n coarse steps replayed from captured hipGraphs. */
  void advance(int n) {
    Check(octlb_dev_advance_graph(dev_, n));
    for (int k = 0; k < n; ++k) Count(0);
  }

  int max_level() const { return max_level_; }
  const TimeLoopCounters& counters() const { return counters_; }
  /** Resets the counters and the step time to 0 (time_loop.h:166-171). */
  void reset_counters() {
    std::fill(counters_.collide.begin(), counters_.collide.end(), 0);
    std::fill(counters_.stream.begin(), counters_.stream.end(), 0);
    counters_.coupler_calls.clear();
    Check(octlb_dev_set_step(dev_, 0));
  }
  void set_snapshot_post_collide(bool on) {
    Check(octlb_dev_set_snapshot(dev_, on ? 1 : 0));
    snapshot_ = on;
  }
  bool snapshot_post_collide() const { return snapshot_; }
  /** ConstRhoBGK average of the previous collide (time_loop.h:88). */
  double average_rho() const {
    double r = 1.0;
    Check(octlb_dev_average_rho(dev_, &r));
    return r;
  }

  /** MomentumExchangeDrag::force_on_fluid_if with the obstacle predicate
   *  "neighbour centre inside bbox" (drag_postprocessor.h:66-115). */
  std::array<double, 3> drag(const BoundingBox& bb) const {
    const double box[6] = {bb.x_min, bb.y_min, bb.z_min, bb.x_max, bb.y_max, bb.z_max};
    std::array<double, 3> f{};
    Check(octlb_dev_drag(dev_, box, f.data()));
    return f;
  }
  /** Sum of rho over kBulk cells (cylinder3d_case.h:601-622). */
  double mass() const {
    double m = 0.0;
    Check(octlb_dev_mass(dev_, &m));
    return m;
  }
  /** Copies the device populations back into the host lattices' halo-
   *  inclusive storage, for host-side observers (VTK, drag, probes). */
  void sync_to_host() {
    for (int b = 0; b < nb_; ++b) Check(octlb_dev_download_block(dev_, b, blocks_[b].populations_at_halo(0, 0, 0)));
  }
  /** Re-reads the host lattices (after a host-side edit of populations). */
  void sync_to_device() {
    for (int b = 0; b < nb_; ++b) Check(octlb_dev_upload_block(dev_, b, blocks_[b].populations_at_halo(0, 0, 0)));
  }
  octlb_dev* device() const { return dev_; }

 private:
  // advance(level) recursion's counter updates (time_loop.h:201-240)
  void Count(int level) {
    ++counters_.collide[static_cast<std::size_t>(level)];
    ++counters_.stream[static_cast<std::size_t>(level)];
    if (level >= max_level_) return;
    counters_.coupler_calls.push_back({CouplerPhase::kHalfTime, level});
    Count(level + 1);
    counters_.coupler_calls.push_back({CouplerPhase::kFullTime, level});
    Count(level + 1);
    counters_.coupler_calls.push_back({CouplerPhase::kRestrict, level});
  }

  void FlattenForest(const OctreeForest& forest) {
    level_.resize(static_cast<std::size_t>(nb_));
    bounds_.resize(6 * static_cast<std::size_t>(nb_));
    for (int b = 0; b < nb_; ++b) {
      level_[b] = forest.quadrant_level(static_cast<OctantId>(b));
      const BoundingBox bb = forest.quadrant_bounds(static_cast<OctantId>(b));
      const double v[6] = {bb.x_min, bb.y_min, bb.z_min, bb.x_max, bb.y_max, bb.z_max};
      std::copy(v, v + 6, &bounds_[6 * static_cast<std::size_t>(b)]);
    }
    // integer block coordinates in units of each block's own width
    double lo[3] = {bounds_[0], bounds_[1], bounds_[2]}, hi[3] = {bounds_[3], bounds_[4], bounds_[5]};
    for (int b = 0; b < nb_; ++b)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], bounds_[6 * b + k]);
        hi[k] = std::max(hi[k], bounds_[6 * b + 3 + k]);
      }
    using Key = std::tuple<int, int64_t, int64_t, int64_t>;
    std::map<Key, int> at;
    std::vector<std::array<int64_t, 3>> ix(static_cast<std::size_t>(nb_));
    for (int b = 0; b < nb_; ++b) {
      const double w = bounds_[6 * b + 3] - bounds_[6 * b];
      for (int k = 0; k < 3; ++k) ix[b][k] = std::llround((bounds_[6 * b + k] - lo[k]) / w);
      at[Key{level_[b], ix[b][0], ix[b][1], ix[b][2]}] = b;
    }
    // FaceDir order: x-, x+, y-, y+, z-, z+ (types.h:15-22)
    for (int b = 0; b < nb_; ++b)
      for (int d = 0; d < 6; ++d) {
        const int axis = d / 2, side = d & 1;
        int64_t c[3] = {ix[b][0], ix[b][1], ix[b][2]};
        c[axis] += side ? 1 : -1;
        const auto it = at.find(Key{level_[b], c[0], c[1], c[2]});
        if (it != at.end()) same_.insert(same_.end(), {b, d, it->second});
        const double face = bounds_[6 * b + axis + (side ? 3 : 0)];
        const double edge = side ? hi[axis] : lo[axis];
        if (std::fabs(face - edge) <= 1e-12 * std::max(1.0, std::fabs(edge))) geo_tree_.insert(geo_tree_.end(), {b, d});
      }
  }

  // CoarseFineFace list (face_pair_list.h:22-34) rebuilt from the coupling
  // plan, in the order the plan's faces first appear
  void FlattenCoupling(const OctreeForest& forest, const LevelCoupler& coupler) {
    std::map<std::pair<int, int>, std::size_t> face_of;
    for (const CouplingPoint& pt : coupler.coupling_plan()) {
      const std::pair<int, int> key{static_cast<int>(pt.coarse_id), static_cast<int>(pt.normal)};
      auto it = face_of.find(key);
      if (it == face_of.end()) {
        octlb_cf_face f{};
        f.coarse_id = pt.coarse_id;
        for (int s = 0; s < 4; ++s) f.fine_ids[s] = -1;
        f.normal = static_cast<int32_t>(pt.normal);
        f.coarse_level = pt.coarse_level;
        f.fine_level = pt.fine_level;
        const BoundingBox cb = forest.quadrant_bounds(pt.coarse_id);
        const double v[6] = {cb.x_min, cb.y_min, cb.z_min, cb.x_max, cb.y_max, cb.z_max};
        std::copy(v, v + 6, f.coarse_bounds);
        it = face_of.emplace(key, cf_.size()).first;
        cf_.push_back(f);
      }
      octlb_cf_face& f = cf_[it->second];
      if (pt.fine_slot < 0 || pt.fine_slot > 3) throw std::invalid_argument("gpu::TimeLoop: coupling point fine slot");
      if (f.fine_ids[pt.fine_slot] < 0) {
        f.fine_ids[pt.fine_slot] = pt.fine_id;
        const BoundingBox fb = forest.quadrant_bounds(pt.fine_id);
        const double v[6] = {fb.x_min, fb.y_min, fb.z_min, fb.x_max, fb.y_max, fb.z_max};
        std::copy(v, v + 6, f.fine_bounds[pt.fine_slot]);
      }
    }
    for (const octlb_cf_face& f : cf_)
      for (int s = 0; s < 4; ++s)
        if (f.fine_ids[s] < 0) throw std::invalid_argument("gpu::TimeLoop: coarse-fine face with a missing fine slot");
  }

  octlb_params FlattenHandler(DomainBoundaryHandler& h, double omega, bool const_rho, ConstRhoStatsScope scope,
                              int precision) {
    octlb_params p{};
    p.omega = omega;
    p.precision = precision;
    p.const_rho = const_rho ? 1 : 0;
    p.stats_scope = scope == ConstRhoStatsScope::kFluidOnly ? 1 : 0;
    if (auto* c = dynamic_cast<ConcreteDomainBoundaryHandler*>(&h)) {
      if (c->padding_mode() != OverlapPaddingMode::kHybrid)
        throw std::invalid_argument("gpu::TimeLoop: the device path runs the kHybrid overlap padding mode");
      p.handler = OCTLB_HANDLER_CONCRETE;
      for (const TreeBoundaryFace& f : c->tree_boundary_faces())
        tree_.insert(tree_.end(), {static_cast<int32_t>(f.octant_id), static_cast<int32_t>(f.face_dir)});
      for (const DomainBcSpec& s : c->specs()) {
        octlb_bc_spec o{};
        o.face = static_cast<int32_t>(s.face);
        o.type = static_cast<int32_t>(s.type);
        for (int d = 0; d < 3; ++d) o.u_wall[d] = s.u_wall[d];
        o.rho_target = s.rho_target;
        o.inlet = OCTLB_INLET_NONE;
        if (s.inlet_field) {
          auto* dev_field = dynamic_cast<const DeviceInletField*>(s.inlet_field.get());
          if (!dev_field)
            throw std::invalid_argument("gpu::TimeLoop: inlet field has no device form (use gpu:: inlet profiles)");
          dev_field->describe(&o);
        }
        specs_.push_back(o);
      }
    } else if (dynamic_cast<NoOpDomainBoundaryHandler*>(&h)) {
      p.handler = OCTLB_HANDLER_NOOP;
      tree_ = geo_tree_;
    } else {
      throw std::invalid_argument(
          "gpu::TimeLoop: domain boundary handler unknown to the device path "
          "(use gpu::ConcreteDomainBoundaryHandler or NoOpDomainBoundaryHandler)");
    }
    // one stream mode for all blocks: the per-cell path where blocks carry
    // Bouzidi links (block_lattice.cpp:549-553)
    int with_links = 0;
    for (int b = 0; b < nb_; ++b) with_links += blocks_[b].bouzidi_links() != nullptr ? 1 : 0;
    if (with_links != 0 && with_links != nb_)
      throw std::invalid_argument("gpu::TimeLoop: either every block or no block carries Bouzidi links");
    p.bouzidi = with_links ? 1 : 0;
    p.snapshot = 0;
    p.num_specs = static_cast<int32_t>(specs_.size());
    p.specs = specs_.empty() ? nullptr : specs_.data();
    p.fast_stats = 0;
    return p;
  }

  void UploadBlocks() {
    const int nc = n_ * n_ * n_;
    std::vector<uint8_t> kinds(static_cast<std::size_t>(nb_) * nc);
    std::vector<int32_t> link;
    std::vector<double> q;
    for (int b = 0; b < nb_; ++b) {
      const BouzidiLinkData* bl = blocks_[b].bouzidi_links();
      for (int x = 0; x < n_; ++x)
        for (int y = 0; y < n_; ++y)
          for (int z = 0; z < n_; ++z) {
            const int c = (x * n_ + y) * n_ + z;
            kinds[static_cast<std::size_t>(b) * nc + c] = static_cast<uint8_t>(blocks_[b].bc_kind(x, y, z));
            if (!bl) continue;
            for (int i = 0; i < TimeLoopLattice::kQ; ++i) {
              const double qf = bl->q_frac(static_cast<OctantId>(b), x, y, z, i);
              if (qf == 1.0) continue;
              link.insert(link.end(), {b, c, i});
              q.push_back(qf);
            }
          }
    }
    Check(octlb_dev_set_bc_kinds(dev_, kinds.data()));
    if (!q.empty()) Check(octlb_dev_set_links(dev_, static_cast<int64_t>(q.size()), link.data(), q.data()));
    sync_to_device();
  }

  BlockCollection<TimeLoopLattice>& blocks_;
  octlb_dev* dev_ = nullptr;
  int nb_ = 0, n_ = 0, halo_ = 0, max_level_ = 0;
  bool snapshot_ = false;
  std::vector<int32_t> level_;
  std::vector<double> bounds_;
  std::vector<int32_t> same_, tree_, geo_tree_;
  std::vector<octlb_cf_face> cf_;
  std::vector<octlb_bc_spec> specs_;
  TimeLoopCounters counters_;
};

}  // namespace gpu
}  // namespace octlb

#endif  // OCTLB_GPU_TIME_LOOP_H_
