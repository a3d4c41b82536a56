// ljg_checkpoint.cpp -- md-flexible's checkpoint format (host side).
//
// A checkpoint is a .pvtu index, <dir>/<name>_Particles_<iteration>.pvtu, listing one piece per rank,
// <dir>/data/<name>_Particles_<rank>_<iteration>.vtu, with the iteration zero-padded to a fixed width
// (ParallelVtkWriter.cpp:60-200, :309-358, generateFilename :438-442).  A piece holds, in this order,
// the point arrays velocities, forces, typeIds, ids and then the positions, as ASCII text.
//
// Writing: the same layout, with 17 significant digits so FP64 state survives a round trip (md-flexible
// itself writes the stream's default 6 digits, raised only near the upper box bound).  One piece.
// Reading: as MDFlexConfig::loadParticlesFromCheckpoint with all pieces loaded by one rank
// (MDFlexConfig.cpp:648-670): the pieces are counted by the word "Piece" in the index, each piece's
// file name is rebuilt from the index name (scenario = up to the last '_', iteration = after it), and
// the arrays are found by name word by word (MDFlexConfig.cpp:33-183).
#include <sys/stat.h>

#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ljgpu.h"

namespace {

bool make_dir(const std::string &path) {
  if (mkdir(path.c_str(), 0755) == 0) return true;
  if (errno != EEXIST) return false;
  struct stat st {};
  return stat(path.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

std::string padded(int64_t value, int32_t width) {
  std::ostringstream s;
  s << std::setfill('0') << std::setw(width) << value;
  return s.str();
}

// Advance the stream past the first occurrence of `word` delimited by the separators md-flexible's
// reader uses; false at end of file.
bool find_word(std::istream &in, const std::string &word) {
  static const char seps[] = {' ', '/', '.', ',', '?', '!', '"', '\'', '<', '>', '=', ':', ';', '\n', '\t', '\r'};
  std::string cur;
  for (;;) {
    const int ch = in.get();
    if (ch == EOF) return false;
    const char c = (char)ch;
    if (std::memchr(seps, c, sizeof seps)) {
      if (cur == word) return true;
      cur.clear();
    } else {
      cur += c;
    }
  }
}

struct Piece {
  std::vector<double> v, f, r;
  std::vector<int64_t> type, id;
};

bool read_piece(const std::string &path, Piece &p) {
  std::ifstream in(path);
  if (!in) return false;
  if (!find_word(in, "NumberOfPoints")) return false;
  in.ignore(std::numeric_limits<std::streamsize>::max(), '"');
  int64_t n = 0;
  in >> n;
  if (!in || n <= 0) return false;
  auto block = [&](const char *name, int comps, auto &out) {
    if (!find_word(in, name)) return false;
    in.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
    out.resize((size_t)(n * comps));
    for (auto &val : out) in >> val;
    return (bool)in;
  };
  return block("velocities", 3, p.v) && block("forces", 3, p.f) && block("typeIds", 1, p.type) &&
         block("ids", 1, p.id) && block("positions", 3, p.r);
}

}  // namespace

extern "C" ljg_status ljg_write_checkpoint(const char *dir, const char *name, int64_t iteration, int32_t digits,
                                           int64_t n, const double *x, const double *y, const double *z,
                                           const double *vx, const double *vy, const double *vz, const double *fx,
                                           const double *fy, const double *fz, const int32_t *type,
                                           const int64_t *id, char *pvtu_path, int32_t pvtu_path_len) {
  if (!dir || !name || n <= 0 || iteration < 0 || digits < 1 || digits > 18 || !x || !y || !z || !vx || !vy ||
      !vz || !fx || !fy || !fz || !type || !id)
    return LJG_ERR_ARG;
  const std::string base(dir), data = base + "/data";
  if (!make_dir(base) || !make_dir(data)) return LJG_ERR_ARG;
  const std::string it = padded(iteration, digits);
  const std::string piece_name = std::string(name) + "_Particles_0_" + it + ".vtu";
  const std::string index = base + "/" + name + "_Particles_" + it + ".pvtu";
  {
    std::ofstream f(data + "/" + piece_name, std::ios::out | std::ios::binary);
    if (!f) return LJG_ERR_ARG;
    f << std::setprecision(17);
    f << "<?xml version=\"1.0\" encoding=\"UTF-8\" standalone=\"no\" ?>\n";
    f << "<VTKFile byte_order=\"LittleEndian\" type=\"UnstructuredGrid\" version=\"0.1\">\n";
    f << "  <UnstructuredGrid>\n";
    f << "    <Piece NumberOfCells=\"0\" NumberOfPoints=\"" << n << "\">\n";
    f << "      <PointData>\n";
    auto vec3 = [&](const char *label, const double *a, const double *b, const double *c) {
      f << "        <DataArray Name=\"" << label << "\" NumberOfComponents=\"3\" format=\"ascii\" type=\"Float32\">\n";
      for (int64_t i = 0; i < n; ++i) f << "        " << a[i] << " " << b[i] << " " << c[i] << "\n";
      f << "        </DataArray>\n";
    };
    vec3("velocities", vx, vy, vz);
    vec3("forces", fx, fy, fz);
    f << "        <DataArray Name=\"typeIds\" NumberOfComponents=\"1\" format=\"ascii\" type=\"Int32\">\n";
    for (int64_t i = 0; i < n; ++i) f << "        " << type[i] << "\n";
    f << "        </DataArray>\n";
    f << "        <DataArray Name=\"ids\" NumberOfComponents=\"1\" format=\"ascii\" type=\"Int32\">\n";
    for (int64_t i = 0; i < n; ++i) f << "        " << id[i] << "\n";
    f << "        </DataArray>\n";
    f << "      </PointData>\n";
    f << "      <CellData/>\n";
    f << "      <Points>\n";
    vec3("positions", x, y, z);
    f << "      </Points>\n";
    f << "      <Cells>\n";
    f << "        <DataArray Name=\"types\" NumberOfComponents=\"0\" format=\"ascii\" type=\"Float32\"/>\n";
    f << "      </Cells>\n";
    f << "    </Piece>\n";
    f << "  </UnstructuredGrid>\n";
    f << "</VTKFile>\n";
    if (!f) return LJG_ERR_ARG;
  }
  {
    std::ofstream f(index, std::ios::out | std::ios::binary);
    if (!f) return LJG_ERR_ARG;
    f << "<?xml version=\"1.0\" encoding=\"UTF-8\" standalone=\"no\" ?>\n";
    f << "<VTKFile byte_order=\"LittleEndian\" type=\"PUnstructuredGrid\" version=\"0.1\">\n";
    f << "  <PUnstructuredGrid GhostLevel=\"0\">\n";
    f << "    <PPointData>\n";
    f << "      <PDataArray Name=\"velocities\" NumberOfComponents=\"3\" format=\"ascii\" type=\"Float32\"/>\n";
    f << "      <PDataArray Name=\"forces\" NumberOfComponents=\"3\" format=\"ascii\" type=\"Float32\"/>\n";
    f << "      <PDataArray Name=\"typeIds\" NumberOfComponents=\"1\" format=\"ascii\" type=\"Int32\"/>\n";
    f << "      <PDataArray Name=\"ids\" NumberOfComponents=\"1\" format=\"ascii\" type=\"Int32\"/>\n";
    f << "    </PPointData>\n";
    f << "    <PCellData/>\n";
    f << "    <PPoints>\n";
    f << "      <PDataArray Name=\"positions\" NumberOfComponents=\"3\" format=\"ascii\" type=\"Float32\"/>\n";
    f << "    </PPoints>\n";
    f << "    <PCells>\n";
    f << "      <PDataArray Name=\"types\" NumberOfComponents=\"0\" format=\"ascii\" type=\"Float32\"/>\n";
    f << "    </PCells>\n";
    f << "    <Piece Source=\"./data/" << piece_name << "\"/>\n";
    f << "  </PUnstructuredGrid>\n";
    f << "</VTKFile>\n";
    if (!f) return LJG_ERR_ARG;
  }
  if (pvtu_path && pvtu_path_len > 0) std::snprintf(pvtu_path, (size_t)pvtu_path_len, "%s", index.c_str());
  return LJG_OK;
}

extern "C" ljg_status ljg_read_checkpoint(const char *pvtu, int64_t *n, double *x, double *y, double *z, double *vx,
                                          double *vy, double *vz, double *fx, double *fy, double *fz,
                                          int32_t *type, int64_t *id) {
  if (!pvtu || !n) return LJG_ERR_ARG;
  const std::string path(pvtu);
  size_t pieces = 0;
  {
    std::ifstream in(path);
    if (!in) return LJG_ERR_ARG;
    while (find_word(in, "Piece")) ++pieces;
  }
  if (pieces == 0) return LJG_ERR_ARG;
  const size_t slash = path.find_last_of('/');
  const std::string folder = slash == std::string::npos ? "." : path.substr(0, slash);
  const std::string file = slash == std::string::npos ? path : path.substr(slash + 1);
  const size_t us = file.find_last_of('_'), dot = file.find_last_of('.');
  if (us == std::string::npos || dot == std::string::npos || dot < us) return LJG_ERR_ARG;
  const std::string scenario = file.substr(0, us), iteration = file.substr(us + 1, dot - us - 1);
  std::vector<Piece> all(pieces);
  int64_t total = 0;
  for (size_t r = 0; r < pieces; ++r) {
    const std::string piece = folder + "/data/" + scenario + "_" + std::to_string(r) + "_" + iteration + ".vtu";
    if (!read_piece(piece, all[r])) return LJG_ERR_ARG;
    total += (int64_t)all[r].id.size();
  }
  if (!x) {
    *n = total;
    return LJG_OK;
  }
  if (*n < total || !y || !z || !vx || !vy || !vz || !fx || !fy || !fz || !type || !id) return LJG_ERR_ARG;
  int64_t k = 0;
  for (const Piece &p : all) {
    for (size_t i = 0; i < p.id.size(); ++i, ++k) {
      x[k] = p.r[3 * i]; y[k] = p.r[3 * i + 1]; z[k] = p.r[3 * i + 2];
      vx[k] = p.v[3 * i]; vy[k] = p.v[3 * i + 1]; vz[k] = p.v[3 * i + 2];
      fx[k] = p.f[3 * i]; fy[k] = p.f[3 * i + 1]; fz[k] = p.f[3 * i + 2];
      type[k] = (int32_t)p.type[i];
      id[k] = p.id[i];
    }
  }
  *n = total;
  return LJG_OK;
}
