// Test-only driver of the reference's own VTK writers (compiled from the
// reference sources by oracle/Makefile into oracle/_ref/ref_vtk; never part
// of the product).  Modes:
//   ref_vtk vts <input.bin> <out.vts>   WriteStructuredGridVts on the block
//                                       described by input.bin
//   ref_vtk vtm <dir> <base> <iT> <rel>...   WriteVtmIndex
//   ref_vtk pvd <pvd_path> <iT> <vtm_rel>    AppendPvdTimestep
// input.bin: int32 nx, ny, nz; double bounds[6]; int32 nfields; per field
// int32 name length, name bytes, int32 components, then nx*ny*nz*components
// doubles in VTK cell order (x fastest).
#include <mpi.h>

#include <array>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "src/solver/io/vtk_writer/parallel_vtk_index.h"
#include "src/solver/io/vtk_writer/structured_grid_writer.h"

namespace {

template <typename V>
V Read(std::ifstream& in) {
  V v{};
  in.read(reinterpret_cast<char*>(&v), sizeof(V));
  if (!in) {
    std::fprintf(stderr, "ref_vtk: short input\n");
    std::exit(2);
  }
  return v;
}

int Vts(const char* input, const char* out) {
  std::ifstream in(input, std::ios::binary);
  octlb::VtkBlockMeta meta;
  meta.nx = Read<int32_t>(in);
  meta.ny = Read<int32_t>(in);
  meta.nz = Read<int32_t>(in);
  meta.bounds.x_min = Read<double>(in);
  meta.bounds.y_min = Read<double>(in);
  meta.bounds.z_min = Read<double>(in);
  meta.bounds.x_max = Read<double>(in);
  meta.bounds.y_max = Read<double>(in);
  meta.bounds.z_max = Read<double>(in);
  const int nf = Read<int32_t>(in);
  const size_t ncell = static_cast<size_t>(meta.nx) * meta.ny * meta.nz;
  std::vector<std::string> names(nf);
  std::vector<int> comps(nf);
  std::vector<std::vector<double>> data(nf);
  for (int f = 0; f < nf; ++f) {
    names[f].resize(Read<int32_t>(in));
    in.read(names[f].data(), static_cast<std::streamsize>(names[f].size()));
    comps[f] = Read<int32_t>(in);
    data[f].resize(ncell * comps[f]);
    for (double& v : data[f]) v = Read<double>(in);
  }
  std::vector<octlb::VtkCellFieldView> views;
  for (int f = 0; f < nf; ++f) {
    const int nx = meta.nx, ny = meta.ny, c = comps[f];
    const std::vector<double>* d = &data[f];
    views.push_back({names[f], c, [=](int i, int j, int k, double* o) {
                       const size_t base = ((static_cast<size_t>(k) * ny + j) * nx + i) * c;
                       for (int q = 0; q < c; ++q) o[q] = (*d)[base + q];
                     }});
  }
  octlb::WriteStructuredGridVts(out, meta, std::span<const octlb::VtkCellFieldView>(views.data(), views.size()));
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  MPI_Init(&argc, &argv);
  int rc = 1;
  const std::string mode = argc > 1 ? argv[1] : "";
  if (mode == "vts" && argc == 4) {
    rc = Vts(argv[2], argv[3]);
  } else if (mode == "vtm" && argc >= 5) {
    std::vector<std::string> rel(argv + 5, argv + argc);
    octlb::WriteVtmIndex(MPI_COMM_WORLD, std::atoi(argv[4]), argv[2], argv[3], rel);
    rc = 0;
  } else if (mode == "pvd" && argc == 5) {
    octlb::AppendPvdTimestep(std::atoi(argv[3]), argv[2], argv[4]);
    rc = 0;
  } else {
    std::fprintf(stderr, "usage: ref_vtk vts|vtm|pvd ...\n");
  }
  MPI_Finalize();
  return rc;
}
