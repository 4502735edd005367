// ref_harness.cpp -- drives the REFERENCE's own per-block solver code
// (BlockLattice, BcDispatcher via ConcreteDomainBoundaryHandler, Skordos FD,
// Bouzidi pull, MomentumExchangeDrag), compiled from /root/reference by
// oracle/Makefile into oracle/_ref/ref_harness.  Test infrastructure only.
//
// It replays TimeLoop::advance for a single-level forest
// (time_loop.h:201-246 with max_level = 0): collide per handler, same-rank
// halo exchange by BlockLattice::pack_*/unpack_* in GhostSchedule's order
// (ghost_schedule.h:23-119, 121-255), snapshot, stream, post-stream BC,
// ConstRho statistics lag.  Multi-level runs need LevelCoupler, whose
// constructor needs a p4est-backed FacePairList/OctreeForest; those are not
// buildable here, so the harness rejects max_level > 0.
//
// Usage: ref_harness <scenario.bin> <out.bin>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "src/solver/lbm/bc_dispatcher.h"
#include "src/solver/lbm/block_lattice.h"
#include "src/solver/lbm/bouzidi_link_data.h"
#include "src/solver/lbm/domain_boundary_handler.h"
#include "src/solver/lbm/drag_postprocessor.h"
#include "src/solver/lbm/boundary/inlet_velocity_field.h"

#include "../include/octlb_gpu.h"

using namespace octlb;
using Lat = BlockLattice<double, olb::descriptors::D3Q19<>>;

namespace {

struct Reader {
  FILE* f;
  template <typename T>
  void read(T* p, size_t count) {
    if (count && std::fread(p, sizeof(T), count, f) != count) throw std::runtime_error("short read");
  }
  template <typename T>
  T get() {
    T v;
    read(&v, 1);
    return v;
  }
};

std::shared_ptr<boundary::InletVelocityField> MakeInlet(const octlb_bc_spec& s) {
  switch (s.inlet) {
    case OCTLB_INLET_UNIFORM:
      return std::make_shared<boundary::UniformInletProfile>(
          std::array<double, 3>{s.inlet_u[0], s.inlet_u[1], s.inlet_u[2]}, s.ramp_end_t);
    case OCTLB_INLET_POISEUILLE:
      return std::make_shared<boundary::PoiseuilleInletProfile>(
          static_cast<FaceDir>(s.inlet_face), s.n_inplane0, s.n_inplane1, s.u_peak,
          std::array<double, 3>{s.inlet_u[0], s.inlet_u[1], s.inlet_u[2]}, s.ramp_end_t);
    case OCTLB_INLET_CHANNEL:
      return std::make_shared<boundary::ChannelPoiseuilleInletProfile>(
          static_cast<FaceDir>(s.inlet_face), s.lo0, s.hi0, s.lo1, s.hi1, s.u_peak,
          std::array<double, 3>{s.inlet_u[0], s.inlet_u[1], s.inlet_u[2]}, s.ramp_end_t);
    default:
      return nullptr;
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s scenario.bin out.bin\n", argv[0]);
    return 2;
  }
  try {
    Reader r{std::fopen(argv[1], "rb")};
    if (!r.f) throw std::runtime_error("cannot open scenario");
    const int32_t magic = r.get<int32_t>();
    if (magic != 0x4f43544c) throw std::runtime_error("bad magic");
    const int n = r.get<int32_t>(), h = r.get<int32_t>(), nb = r.get<int32_t>();
    const int handler = r.get<int32_t>(), const_rho = r.get<int32_t>();
    const int stats_scope = r.get<int32_t>(), bouzidi = r.get<int32_t>();
    const int snapshot = r.get<int32_t>(), steps = r.get<int32_t>();
    const int num_specs = r.get<int32_t>(), nsame = r.get<int32_t>();
    const int num_links = r.get<int32_t>(), num_tb = r.get<int32_t>();
    const int t0 = r.get<int32_t>();
    const int max_level = r.get<int32_t>();
    const double omega = r.get<double>();
    if (max_level != 0) throw std::runtime_error("harness drives single-level forests only");
    std::vector<octlb_bc_spec> raw(num_specs);
    r.read(raw.data(), raw.size());
    std::vector<double> bounds(static_cast<size_t>(nb) * 6);
    r.read(bounds.data(), bounds.size());
    std::vector<int32_t> same(static_cast<size_t>(nsame) * 3);
    r.read(same.data(), same.size());
    std::vector<int32_t> tbv(static_cast<size_t>(num_tb) * 2);
    r.read(tbv.data(), tbv.size());
    const int ncell = n * n * n;
    std::vector<uint8_t> kinds(static_cast<size_t>(nb) * ncell);
    r.read(kinds.data(), kinds.size());
    std::vector<int32_t> links(static_cast<size_t>(num_links) * 3);
    r.read(links.data(), links.size());
    std::vector<double> qv(num_links);
    r.read(qv.data(), qv.size());
    const int NP = n + 2 * h;
    const size_t nslot = static_cast<size_t>(NP) * NP * NP;
    std::vector<double> pops(static_cast<size_t>(nb) * nslot * 19);
    r.read(pops.data(), pops.size());
    double bbox[6];
    r.read(bbox, 6);
    std::fclose(r.f);

    std::vector<DomainBcSpec> specs;
    for (const auto& s : raw) {
      DomainBcSpec d;
      d.face = static_cast<FaceDir>(s.face);
      d.type = static_cast<DomainBcType>(s.type);
      d.u_wall = {s.u_wall[0], s.u_wall[1], s.u_wall[2]};
      d.rho_target = s.rho_target;
      d.inlet_field = MakeInlet(s);
      specs.push_back(d);
    }

    BouzidiLinkData link_data(nb, n, n, n, 19);
    for (int i = 0; i < num_links; ++i) {
      const int b = links[3 * i], c = links[3 * i + 1], p = links[3 * i + 2];
      link_data.set_q_frac(b, c / (n * n), (c / n) % n, c % n, p, qv[i]);
    }

    BlockCollection<Lat> blocks(nb, [&](OctantId id) {
      Lat lat(n, n, n, h);
      lat.set_octant_id(id);
      const double* bb = &bounds[static_cast<size_t>(id) * 6];
      lat.set_phys_origin({bb[0], bb[1], bb[2]}, (bb[3] - bb[0]) / static_cast<double>(n));
      for (int ix = 0; ix < n; ++ix)
        for (int iy = 0; iy < n; ++iy)
          for (int iz = 0; iz < n; ++iz)
            lat.set_bc_kind(ix, iy, iz,
                            static_cast<BcKind>(kinds[static_cast<size_t>(id) * ncell +
                                                      (ix * n + iy) * n + iz]));
      for (int hx = 0; hx < NP; ++hx)
        for (int hy = 0; hy < NP; ++hy)
          for (int hz = 0; hz < NP; ++hz) {
            double* p = lat.populations_at_halo(hx, hy, hz);
            std::memcpy(p, &pops[(static_cast<size_t>(id) * nslot + (hx * NP + hy) * NP + hz) * 19],
                        sizeof(double) * 19);
          }
      if (bouzidi) lat.set_bouzidi_links(&link_data);
      return lat;
    });

    std::vector<TreeBoundaryFace> tree_faces;
    for (int i = 0; i < num_tb; ++i)
      tree_faces.push_back({tbv[2 * i], static_cast<FaceDir>(tbv[2 * i + 1])});
    std::unique_ptr<DomainBoundaryHandler> bc;
    if (handler == OCTLB_HANDLER_CONCRETE) {
      bc = std::make_unique<ConcreteDomainBoundaryHandler>(blocks, tree_faces, specs, n, n, n,
                                                           omega);
    } else {
      bc = std::make_unique<NoOpDomainBoundaryHandler>();
    }

    // GhostSchedule same-rank plan (ghost_schedule.h:40-91).
    std::vector<OctantId> face_nbr(static_cast<size_t>(nb) * 6, -1);
    for (int e = 0; e < nsame; ++e) face_nbr[same[3 * e] * 6 + same[3 * e + 1]] = same[3 * e + 2];
    struct Edge {
      int a, d1, d2, dd;
    };
    std::vector<Edge> edges;
    for (int a = 0; a < nb; ++a)
      for (int d1 = 0; d1 < 6; ++d1)
        for (int d2 = 0; d2 < 6; ++d2) {
          if (d1 / 2 == d2 / 2) continue;
          const int b = face_nbr[a * 6 + d1];
          if (b < 0) continue;
          const int dd = face_nbr[b * 6 + d2];
          if (dd < 0) continue;
          edges.push_back({a, d1, d2, dd});
        }
    std::vector<double> buf(static_cast<size_t>(n) * n * 19);
    auto exchange = [&]() {
      for (int e = 0; e < nsame; ++e) {
        const FaceDir dir = static_cast<FaceDir>(same[3 * e + 1]);
        const int cnt = Lat::face_buffer_count(n, n, n, dir);
        blocks[same[3 * e]].pack_face(dir, buf.data(), cnt);
        blocks[same[3 * e + 2]].unpack_face(static_cast<FaceDir>(static_cast<int>(dir) ^ 1),
                                            buf.data(), cnt);
      }
      for (const Edge& e : edges) {
        const FaceDir d1 = static_cast<FaceDir>(e.d1), d2 = static_cast<FaceDir>(e.d2);
        const int cnt = Lat::edge_buffer_count(n, n, n, d1, d2);
        blocks[e.a].pack_edge(d1, d2, buf.data(), cnt);
        blocks[e.dd].unpack_edge(static_cast<FaceDir>(e.d1 ^ 1), static_cast<FaceDir>(e.d2 ^ 1),
                                 buf.data(), cnt);
      }
    };

    // TimeLoop::advance(0) with max_level = 0 (time_loop.h:201-246).
    double average_rho = 1.0;
    const bool openlb_order = bc->collide_boundary_before_bulk();
    for (int step = 0; step < steps; ++step) {
      bc->set_time(static_cast<double>(t0 + step));
      CollideRhoStats stats;
      CollideRhoStats* sp = const_rho ? &stats : nullptr;
      for (int id = 0; id < nb; ++id) {
        if (openlb_order) {
          bc->collide_interleaved_with(blocks[id], sp, average_rho, const_rho != 0);
        } else if (const_rho) {
          blocks[id].collide_const_rho(omega, average_rho, sp);
        } else {
          blocks[id].collide(omega);
        }
      }
      if (!openlb_order) {
        bc->apply(sp != nullptr && stats_scope == 0 ? sp : nullptr, average_rho, const_rho != 0);
      }
      exchange();
      if (snapshot)
        for (int id = 0; id < nb; ++id) blocks[id].take_post_collide_snapshot();
      for (int id = 0; id < nb; ++id) blocks[id].stream();
      bc->apply_post_stream();
      if (sp != nullptr) average_rho = stats.average();
    }

    std::array<double, 3> drag{0.0, 0.0, 0.0};
    double mass = 0.0;
    for (int id = 0; id < nb; ++id) {
      const Lat& lat = blocks[id];
      const std::array<double, 3> org = lat.phys_origin();
      const double cw = lat.phys_cell_width();
      auto in_box = [&](int i, int j, int k) {
        const double x = org[0] + (i + 0.5) * cw;
        const double y = org[1] + (j + 0.5) * cw;
        const double z = org[2] + (k + 0.5) * cw;
        return x >= bbox[0] && x <= bbox[3] && y >= bbox[1] && y <= bbox[4] && z >= bbox[2] &&
               z <= bbox[5];
      };
      if (h == 1) {  // MomentumExchangeDrag assumes a unit halo (drag_postprocessor.h:86-88)
        const std::array<double, 3> f = MomentumExchangeDrag(lat).force_on_fluid_if(in_box);
        drag[0] += f[0];
        drag[1] += f[1];
        drag[2] += f[2];
      }
      for (int iz = 0; iz < n; ++iz)
        for (int iy = 0; iy < n; ++iy)
          for (int ix = 0; ix < n; ++ix) {
            if (lat.bc_kind(ix, iy, iz) != BcKind::kBulk) continue;
            double rho = 0.0, u[3] = {};
            lat.get(ix, iy, iz).computeRhoU(rho, u);
            mass += rho;
          }
    }

    FILE* o = std::fopen(argv[2], "wb");
    if (!o) throw std::runtime_error("cannot open output");
    for (int id = 0; id < nb; ++id)
      for (int hx = 0; hx < NP; ++hx)
        for (int hy = 0; hy < NP; ++hy)
          for (int hz = 0; hz < NP; ++hz)
            std::fwrite(blocks[id].populations_at_halo(hx, hy, hz), sizeof(double), 19, o);
    const double tail[5] = {average_rho, mass, drag[0], drag[1], drag[2]};
    std::fwrite(tail, sizeof(double), 5, o);
    std::fclose(o);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "ref_harness: %s\n", e.what());
    return 1;
  }
  return 0;
}
