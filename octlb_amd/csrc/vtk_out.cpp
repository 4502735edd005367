// VTK output of device fields in the reference's file formats, so that
// ParaView workflows built on the reference keep working:
//   .vts  one StructuredGrid per block, CellData Float64 arrays as VTK XML
//         binary (base64 of a UInt32 byte count followed by the raw
//         little-endian values), then the block's (n+1)^3 corner points
//         (WriteStructuredGridVts, structured_grid_writer.cpp:61-117)
//   .vtm  vtkMultiBlockDataSet listing the block files (WriteVtmIndex,
//         parallel_vtk_index.cpp:132-158)
//   .pvd  Collection with one DataSet per written time step, appended in
//         place (AppendPvdTimestep, parallel_vtk_index.cpp:160-186)
// The text around the arrays is the reference's, byte for byte; the point
// coordinates use its interpolation expression so they round the same way.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "octlb_gpu.h"
#include "octlb_internal.h"

namespace octlb {
namespace {

// RFC 4648 base64 with '=' padding
std::string Base64(const unsigned char* p, size_t n) {
  static const char kAlphabet[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  std::string s;
  s.reserve((n + 2) / 3 * 4);
  size_t i = 0;
  for (; i + 3 <= n; i += 3) {
    const uint32_t v = (uint32_t(p[i]) << 16) | (uint32_t(p[i + 1]) << 8) | uint32_t(p[i + 2]);
    s += kAlphabet[v >> 18];
    s += kAlphabet[(v >> 12) & 63];
    s += kAlphabet[(v >> 6) & 63];
    s += kAlphabet[v & 63];
  }
  if (i < n) {
    const bool two = i + 1 < n;
    const uint32_t v = (uint32_t(p[i]) << 16) | (two ? uint32_t(p[i + 1]) << 8 : 0u);
    s += kAlphabet[v >> 18];
    s += kAlphabet[(v >> 12) & 63];
    s += two ? kAlphabet[(v >> 6) & 63] : '=';
    s += '=';
  }
  return s;
}

// VTK XML binary array: UInt32 header (payload bytes) + payload, one base64 run
std::string BinaryArray(const double* v, size_t count) {
  const uint32_t bytes = static_cast<uint32_t>(count * sizeof(double));
  std::vector<unsigned char> raw(sizeof(uint32_t) + count * sizeof(double));
  std::memcpy(raw.data(), &bytes, sizeof(bytes));
  if (count) std::memcpy(raw.data() + sizeof(uint32_t), v, count * sizeof(double));
  return Base64(raw.data(), raw.size());
}

// corner coordinate i of n along [a, b] (structured_grid_writer.cpp:14-19)
double Corner(double a, double b, int i, int n) {
  if (n <= 0) return a;
  return a + (b - a) * static_cast<double>(i) / static_cast<double>(n);
}

}  // namespace

void WriteVts(const std::string& path, int nx, int ny, int nz, const double* bounds, int nfields,
              const char* const* names, const int32_t* comps, const double* const* data) {
  if (nx <= 0 || ny <= 0 || nz <= 0) throw std::invalid_argument("vts: nx, ny, nz must be positive");
  const int npts = (nx + 1) * (ny + 1) * (nz + 1), ncells = nx * ny * nz;
  std::vector<double> pts(static_cast<size_t>(npts) * 3);
  size_t w = 0;
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        pts[w++] = Corner(bounds[0], bounds[3], i, nx);
        pts[w++] = Corner(bounds[1], bounds[4], j, ny);
        pts[w++] = Corner(bounds[2], bounds[5], k, nz);
      }
  std::ostringstream o;
  o << "<?xml version=\"1.0\"?>\n"
    << "<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt32\">\n"
    << "  <StructuredGrid WholeExtent=\"0 " << nx << " 0 " << ny << " 0 " << nz << "\">\n"
    << "    <Piece Extent=\"0 " << nx << " 0 " << ny << " 0 " << nz << "\" NumberOfPoints=\"" << npts
    << "\" NumberOfCells=\"" << ncells << "\">\n"
    << "      <PointData>\n"
    << "      </PointData>\n"
    << "      <CellData>\n";
  for (int f = 0; f < nfields; ++f) {
    o << "        <DataArray type=\"Float64\" Name=\"" << names[f] << "\" NumberOfComponents=\"" << comps[f]
      << "\" format=\"binary\" encoding=\"base64\">\n          "
      << BinaryArray(data[f], static_cast<size_t>(ncells) * comps[f]) << "\n        </DataArray>\n";
  }
  o << "      </CellData>\n"
    << "      <Points>\n"
    << "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"binary\" encoding=\"base64\">\n          "
    << BinaryArray(pts.data(), pts.size()) << "\n        </DataArray>\n"
    << "      </Points>\n"
    << "    </Piece>\n"
    << "  </StructuredGrid>\n"
    << "</VTKFile>\n";
  std::ofstream out(path, std::ios::binary | std::ios::trunc);
  if (!out) throw std::runtime_error("vts: failed to open " + path);
  const std::string s = o.str();
  out.write(s.data(), static_cast<std::streamsize>(s.size()));
  if (!out) throw std::runtime_error("vts: failed to write " + path);
}

void WriteVtm(const std::string& path, int count, const char* const* rel) {
  std::ofstream out(path, std::ios::binary | std::ios::trunc);
  if (!out) throw std::runtime_error("vtm: failed to open " + path);
  out << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"vtkMultiBlockDataSet\" version=\"1.0\" byte_order=\"LittleEndian\">\n"
      << "<vtkMultiBlockDataSet>\n";
  for (int i = 0; i < count; ++i)
    out << "<Block index=\"" << i << "\">\n<DataSet index=\"0\" file=\"" << rel[i] << "\"/>\n</Block>\n";
  out << "</vtkMultiBlockDataSet>\n</VTKFile>\n";
  if (!out) throw std::runtime_error("vtm: failed to write " + path);
}

void AppendPvd(const std::string& path, int iT, const std::string& vtm_rel) {
  static const char kClose[] = "</Collection>\n</VTKFile>\n";  // 25 bytes, rewritten on every append
  std::string body;
  {
    std::ifstream in(path, std::ios::binary);
    if (in) {
      std::ostringstream ss;
      ss << in.rdbuf();
      body = ss.str();
    } else {
      body = std::string("<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\" "
                         "byte_order=\"LittleEndian\">\n<Collection>\n") + kClose;
    }
  }
  const size_t tail = sizeof(kClose) - 1;
  if (body.size() < tail) throw std::runtime_error("pvd: " + path + " is not a collection file");
  body.resize(body.size() - tail);
  std::ostringstream o;
  o << "<DataSet timestep=\"" << iT << "\" group=\"\" part=\"0\" file=\"" << vtm_rel << "\"/>\n" << kClose;
  body += o.str();
  std::ofstream out(path, std::ios::binary | std::ios::trunc);
  if (!out) throw std::runtime_error("pvd: failed to open " + path);
  out.write(body.data(), static_cast<std::streamsize>(body.size()));
  if (!out) throw std::runtime_error("pvd: failed to write " + path);
}

namespace {
template <typename F>
int VtkGuard(F&& f) {
  try {
    f();
  } catch (const std::invalid_argument& e) {
    return SetError(OCTLB_EINVAL, e.what());
  } catch (const std::bad_alloc&) {
    return SetError(OCTLB_ENOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return SetError(OCTLB_EIO, e.what());
  }
  return OCTLB_OK;
}
}  // namespace

}  // namespace octlb

extern "C" {

int octlb_vtk_write_vts(const char* path, int32_t nx, int32_t ny, int32_t nz, const double bounds[6],
                        int32_t num_fields, const char* const* names, const int32_t* components,
                        const double* const* data) {
  return octlb::VtkGuard([&] {
    if (!path || !bounds || num_fields < 0 || (num_fields > 0 && (!names || !components || !data)))
      throw std::invalid_argument("vts: null argument");
    for (int f = 0; f < num_fields; ++f)
      if (!names[f] || !data[f] || components[f] < 1 || components[f] > 3)
        throw std::invalid_argument("vts: fields need a name, data and 1 to 3 components");
    octlb::WriteVts(path, nx, ny, nz, bounds, num_fields, names, components, data);
  });
}

int octlb_vtk_write_vtm(const char* path, int32_t count, const char* const* relative_paths) {
  return octlb::VtkGuard([&] {
    if (!path || count < 0 || (count > 0 && !relative_paths)) throw std::invalid_argument("vtm: null argument");
    for (int i = 0; i < count; ++i)
      if (!relative_paths[i]) throw std::invalid_argument("vtm: null path");
    octlb::WriteVtm(path, count, relative_paths);
  });
}

int octlb_vtk_append_pvd(const char* pvd_path, int32_t iT, const char* vtm_relative_path) {
  return octlb::VtkGuard([&] {
    if (!pvd_path || !vtm_relative_path) throw std::invalid_argument("pvd: null argument");
    octlb::AppendPvd(pvd_path, iT, vtm_relative_path);
  });
}

}  // extern "C"
