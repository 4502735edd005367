// Internal helpers shared by the host and device translation units.
#ifndef OCTLB_AMD_CSRC_OCTLB_INTERNAL_H_
#define OCTLB_AMD_CSRC_OCTLB_INTERNAL_H_

#include <cstdint>
#include <string>
#include <vector>

struct octlb_dev;

namespace octlb {

// Records msg as the calling thread's last error and returns code.
int SetError(int code, const std::string& msg);

// What the discrete adjoint takes from a forward handle: one level of 8^3
// blocks with a one-cell halo (blocked storage), the concrete handler,
// plain BGK and Bouzidi streaming.
struct AdjTopo {
  int nb = 0;
  int prec = 64;                  // forward precision of the handle
  double omega = 0.0;
  std::vector<uint8_t> kinds;     // [nb][512] BcKind bytes
  std::vector<int64_t> link_key;  // (block*512 + cell)*19 + pop of every set link
  std::vector<double> link_q;     // its wall fraction
  std::vector<double> origin_cw;  // [nb][4] origin x, y, z and cell width
  std::vector<int> halo_src;      // [nb][1000] exchanged halo slot -> source block*1000 + slot, or -1
  std::vector<int> level;         // [nb] refinement level (empty: one level)
};

// VTK writers (vtk_out.cpp); throw on bad arguments or file errors.
void WriteVts(const std::string& path, int nx, int ny, int nz, const double* bounds, int nfields,
              const char* const* names, const int32_t* comps, const double* const* data);
void WriteVtm(const std::string& path, int count, const char* const* rel);
void AppendPvd(const std::string& path, int iT, const std::string& vtm_rel);

// Fills t from dev; throws std::invalid_argument when the handle is outside
// what the adjoint differentiates.
void ExportAdjTopo(const struct ::octlb_dev* dev, AdjTopo* t);

}  // namespace octlb

#endif  // OCTLB_AMD_CSRC_OCTLB_INTERNAL_H_
