// What the multi-level adjoint takes from a forward handle beyond AdjTopo:
// the coupling plan, the FD pass tables and the device constants the
// production FD and coupler code reads.  Only the .hip translation units
// include this header.
#ifndef OCTLB_AMD_CSRC_ADJOINT_TOPO_H_
#define OCTLB_AMD_CSRC_ADJOINT_TOPO_H_

#include <vector>

#include "lbm_device.h"
#include "octlb_internal.h"

namespace octlb {

struct AdjMLTopo {
  int max_level = 0;
  std::vector<int> level;                            // [nb]
  std::vector<int> plan;                             // 12 ints per coupling point (CouplingPoint)
  std::vector<std::vector<int>> prolong_active;      // per coarse level: deduplicated points
  std::vector<std::vector<int>> restrict_active;     // per fine level
  double prolong_scale = 1.0, restrict_scale = 1.0;  // (tau - 1/4)/tau, tau/(tau - 1/4)
  std::vector<int> fd_blocks;                        // blocks with FD cells
  std::vector<short> fd_slot, fd_list;               // [fd block][512]
  std::vector<int4> fd_count;                        // [fd block] phase ends
  std::vector<BlockTab> bt;                          // [nb]
  SimConst sc;                                       // the handle's constants (one member)
};

// Fills t (the AdjTopo part, for any number of levels and FD kinds) and m;
// throws std::invalid_argument outside what the multi-level adjoint covers.
void ExportAdjMLTopo(struct ::octlb_dev* dev, AdjTopo* t, AdjMLTopo* m);

// True when the handle has more than one level or FD boundary cells: the
// single-level engine (fused kernels) does not cover it, the multi-level one
// does.
bool AdjNeedsMultiLevel(const struct ::octlb_dev* dev);

}  // namespace octlb

#endif  // OCTLB_AMD_CSRC_ADJOINT_TOPO_H_
