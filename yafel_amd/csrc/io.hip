// Mesh input: the gmsh 2.2 ASCII reader behind the reference's
// Mesh(const std::string &fname) (src/mesh/Mesh.cpp:32-36 ->
// Mesh::parse_gmsh, src/mesh/parse_gmsh.cpp:80-151).  Host code only: the
// arrays it returns feed yafel_hip_mesh_create_cg / yafel_hip_dg_create.
//
// Parsing follows the reference line by line: words split on single spaces
// (runs of spaces give no empty words); "$Nodes" / "$Elements" are followed by
// a count line; a node line is "id x y z" stored at id-1; an element line is
// "id gmsh-type ntags tag... node..." stored at id-1 with 0-based nodes, its
// offset the running count of node entries in file order; every other line
// (the $MeshFormat block, $PhysicalNames, ...) is skipped.  gmsh types map to
// the reference's CellType (parse_gmsh.cpp:156-174); an unknown type becomes
// None.  Beyond the reference: a trailing '\r' is dropped, and a malformed
// line or an id outside the declared count is an error instead of undefined
// behaviour.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace yh;

struct yafel_hip_gmsh_s {
  std::vector<double> nodes;        // 3 per node
  std::vector<int32_t> cell_nodes;  // 0-based
  std::vector<int32_t> cell_offsets;
  std::vector<int32_t> cell_types;  // YAFEL_HIP_CELL_*
  std::vector<int32_t> tag_offsets;
  std::vector<int32_t> tags;
};

namespace {

void split_words(const std::string &s, std::vector<std::string> &w) {
  w.clear();
  std::size_t start = 0;
  while (start < s.size()) {
    std::size_t pos = s.find(' ', start);
    if (pos == std::string::npos) pos = s.size();
    if (pos > start) w.push_back(s.substr(start, pos - start));
    start = pos + 1;
  }
}

long long to_int(const std::string &s, const char *what, long long line) {
  char *end = nullptr;
  errno = 0;
  const long long v = std::strtoll(s.c_str(), &end, 10);
  if (errno || end == s.c_str() || *end)
    fail(YAFEL_HIP_EINVAL, "gmsh_read: bad " + std::string(what) + " '" + s + "' on line " + std::to_string(line));
  return v;
}

double to_real(const std::string &s, long long line) {
  char *end = nullptr;
  errno = 0;
  const double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str() || *end)
    fail(YAFEL_HIP_EINVAL, "gmsh_read: bad coordinate '" + s + "' on line " + std::to_string(line));
  return v;
}

// parse_gmsh.cpp:156-174
int cell_type_of_gmsh(long long id) {
  switch (id) {
    case 1: return YAFEL_HIP_CELL_LINE2;
    case 2: return YAFEL_HIP_CELL_TRI3;
    case 3: return YAFEL_HIP_CELL_QUAD4;
    case 4: return YAFEL_HIP_CELL_TET4;
    case 5: return YAFEL_HIP_CELL_HEX8;
    case 15: return YAFEL_HIP_CELL_POINT1;
    default: return YAFEL_HIP_CELL_NONE;
  }
}

}  // namespace

extern "C" {

int yafel_hip_gmsh_read(const char *path, yafel_hip_gmsh *out) {
  YH_TRY_BEGIN
  REQUIRE(path && out, "gmsh_read: null argument");
  *out = nullptr;
  std::ifstream in(path);
  REQUIRE(in.good(), std::string("gmsh_read: cannot open ") + path);
  auto g = std::make_unique<yafel_hip_gmsh_s>();
  enum { NONE, NODES, ELEMENTS } action = NONE;
  std::string line;
  std::vector<std::string> w;
  long long lineno = 0, nnodes = 0, nelems = 0;
  int64_t offset = 0;
  std::vector<std::vector<int32_t>> cell_tags, cell_conn;
  std::vector<int32_t> cell_off_file;
  auto next_line = [&]() {
    if (!std::getline(in, line)) return false;
    ++lineno;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    return true;
  };
  while (next_line()) {
    split_words(line, w);
    if (w.empty()) continue;
    if (w[0] == "$Nodes") {
      REQUIRE(next_line(), "gmsh_read: missing node count");
      split_words(line, w);
      REQUIRE(!w.empty(), "gmsh_read: missing node count");
      nnodes = to_int(w[0], "node count", lineno);
      REQUIRE(nnodes >= 0 && nnodes < (1LL << 31), "gmsh_read: node count out of range");
      g->nodes.assign((size_t)nnodes * 3, 0.0);
      action = NODES;
      continue;
    }
    if (w[0] == "$EndNodes" || w[0] == "$EndElements") {
      action = NONE;
      continue;
    }
    if (w[0] == "$Elements") {
      REQUIRE(next_line(), "gmsh_read: missing element count");
      split_words(line, w);
      REQUIRE(!w.empty(), "gmsh_read: missing element count");
      nelems = to_int(w[0], "element count", lineno);
      REQUIRE(nelems >= 0 && nelems < (1LL << 31), "gmsh_read: element count out of range");
      g->cell_types.assign((size_t)nelems, YAFEL_HIP_CELL_NONE);
      cell_tags.assign((size_t)nelems, {});
      cell_conn.assign((size_t)nelems, {});
      cell_off_file.assign((size_t)nelems, 0);
      action = ELEMENTS;
      continue;
    }
    if (action == NODES) {
      REQUIRE(w.size() >= 4, "gmsh_read: short node line " + std::to_string(lineno));
      const long long id = to_int(w[0], "node id", lineno);
      REQUIRE(id >= 1 && id <= nnodes, "gmsh_read: node id out of range on line " + std::to_string(lineno));
      for (int i = 0; i < 3; ++i) g->nodes[(size_t)(id - 1) * 3 + i] = to_real(w[i + 1], lineno);
    } else if (action == ELEMENTS) {
      REQUIRE(w.size() >= 3, "gmsh_read: short element line " + std::to_string(lineno));
      const long long id = to_int(w[0], "element id", lineno);
      REQUIRE(id >= 1 && id <= nelems, "gmsh_read: element id out of range on line " + std::to_string(lineno));
      const long long ntags = to_int(w[2], "tag count", lineno);
      REQUIRE(ntags >= 0 && 3 + ntags <= (long long)w.size(),
              "gmsh_read: bad tag count on line " + std::to_string(lineno));
      const size_t e = (size_t)(id - 1);
      g->cell_types[e] = cell_type_of_gmsh(to_int(w[1], "element type", lineno));
      cell_tags[e].clear();
      for (long long i = 3; i < 3 + ntags; ++i) cell_tags[e].push_back((int32_t)to_int(w[i], "tag", lineno));
      cell_conn[e].clear();
      for (size_t i = 3 + (size_t)ntags; i < w.size(); ++i) {
        const long long nd = to_int(w[i], "node", lineno);
        REQUIRE(nd >= 1 && nd <= nnodes, "gmsh_read: element node out of range on line " + std::to_string(lineno));
        cell_conn[e].push_back((int32_t)(nd - 1));
      }
      cell_off_file[e] = (int32_t)offset;
      offset += (int64_t)cell_conn[e].size();
      REQUIRE(offset < (1LL << 31), "gmsh_read: more than 2^31 cell node entries");
    }
  }
  // cells in id order; each id's nodes at its file-order offset (the reference
  // stores cellOffsets_[id-1] = running offset and appends nodes in file order)
  g->cell_nodes.assign((size_t)offset, 0);
  g->cell_offsets.assign((size_t)nelems + 1, (int32_t)offset);
  for (size_t e = 0; e < (size_t)nelems; ++e) {
    g->cell_offsets[e] = cell_off_file[e];
    for (size_t k = 0; k < cell_conn[e].size(); ++k) g->cell_nodes[(size_t)cell_off_file[e] + k] = cell_conn[e][k];
  }
  g->tag_offsets.assign((size_t)nelems + 1, 0);
  for (size_t e = 0; e < (size_t)nelems; ++e) {
    g->tag_offsets[e + 1] = g->tag_offsets[e] + (int32_t)cell_tags[e].size();
    g->tags.insert(g->tags.end(), cell_tags[e].begin(), cell_tags[e].end());
  }
  *out = g.release();
  return YAFEL_HIP_OK;
  YH_TRY_END(nullptr)
}

int yafel_hip_gmsh_sizes(yafel_hip_gmsh g, int64_t *sizes) {
  YH_TRY_BEGIN
  REQUIRE(g && sizes, "gmsh_sizes: null argument");
  sizes[0] = (int64_t)(g->nodes.size() / 3);
  sizes[1] = (int64_t)g->cell_types.size();
  sizes[2] = (int64_t)g->cell_nodes.size();
  sizes[3] = (int64_t)g->tags.size();
  return YAFEL_HIP_OK;
  YH_TRY_END(nullptr)
}

int yafel_hip_gmsh_arrays(yafel_hip_gmsh g, double *nodes, int32_t *cell_nodes, int32_t *cell_offsets,
                          int32_t *cell_types, int32_t *tag_offsets, int32_t *tags) {
  YH_TRY_BEGIN
  REQUIRE(g, "gmsh_arrays: null handle");
  auto put = [](auto *dst, const auto &v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
  };
  put(nodes, g->nodes);
  put(cell_nodes, g->cell_nodes);
  put(cell_offsets, g->cell_offsets);
  put(cell_types, g->cell_types);
  put(tag_offsets, g->tag_offsets);
  put(tags, g->tags);
  return YAFEL_HIP_OK;
  YH_TRY_END(nullptr)
}

int yafel_hip_gmsh_destroy(yafel_hip_gmsh g) {
  delete g;
  return YAFEL_HIP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- output
// VTUBackend (src/output/VTUBackend.cpp:19-230) writing one frame of an
// OutputMesh (src/output/OutputMesh.cpp:11-48): every element of order p is
// drawn as its p^dim local sub-cells (make_tensorproduct_element.cpp:51-95,
// corners (i,j), (i,j+1), (i+1,j+1), (i+1,j) of the x-fastest local node
// grid, and the hexahedral analogue), VTK type 9 (quad) or 12 (hex).  The
// text is the reference's: ostream default formatting of doubles (%g, six
// significant digits), one cell per connectivity line, point data before
// cell data, a cell value repeated on each of its sub-cells.
namespace {

struct VtuOut {
  std::FILE *f = nullptr;
  ~VtuOut() {
    if (f) std::fclose(f);
  }
};

void put_num(std::FILE *f, double v) { std::fprintf(f, "%g", v); }

void write_array(std::FILE *f, const yafel_hip_output_data &d, const std::vector<double> &V, int64_t n_locs) {
  std::vector<int> comps;
  for (int c = 0; c < d.dof_per_loc; ++c)
    if ((d.mask >> c) & 1u) comps.push_back(c);
  const int ncomps = d.data_type == 0 ? 1 : d.data_type == 1 ? 3 : 9;
  std::fprintf(f, "<DataArray type=\"Float32\" format=\"ascii\" Name=\"%s\" NumberOfComponents=\"%d\">\n", d.name,
               ncomps);
  static const char v_sep[3] = {' ', ' ', '\n'};
  for (int64_t n = 0; n < n_locs; ++n) {
    if (d.data_type == 0) {
      put_num(f, V[(size_t)(n * d.dof_per_loc + comps[0])]);
      std::fputc('\n', f);
    } else {
      // VTUBackend.cpp:164-184: vectors (and tensors) are written with the
      // vector separators, padded with zeros to three components
      int vc = 0;
      for (int c : comps) {
        put_num(f, V[(size_t)(n * d.dof_per_loc + c)]);
        std::fputc(v_sep[vc++], f);
      }
      for (int i = vc; i < 3; ++i) {
        std::fputc('0', f);
        std::fputc(v_sep[i], f);
      }
    }
  }
  std::fprintf(f, "</DataArray>\n");
}

}  // namespace

extern "C" int yafel_hip_vtu_write(const char *fname_base, int dim, int p, int64_t n_points, const double *points,
                                   int64_t n_cells, const int32_t *cells, int n_data,
                                   const yafel_hip_output_data *data) {
  YH_TRY_BEGIN
  REQUIRE(fname_base && points && cells, "vtu_write: null argument");
  REQUIRE((dim == 2 || dim == 3) && p >= 1 && p <= 8, "vtu_write: dim must be 2 or 3 and p 1..8");
  REQUIRE(n_points >= 0 && n_cells >= 0 && n_data >= 0 && (n_data == 0 || data), "vtu_write: bad sizes");
  const int n1 = p + 1, nen = dim == 2 ? n1 * n1 : n1 * n1 * n1;
  const int nsub = dim == 2 ? p * p : p * p * p, nv = dim == 2 ? 4 : 8;
  for (int64_t i = 0; i < n_cells * nen; ++i)
    REQUIRE(cells[i] >= 0 && cells[i] < n_points, "vtu_write: cell node out of range");
  // data checks (VTUBackend.cpp:104-131) and the values on the host
  std::vector<std::vector<double>> vals((size_t)n_data);
  for (int k = 0; k < n_data; ++k) {
    const yafel_hip_output_data &d = data[k];
    REQUIRE(d.name && d.data, "vtu_write: data without a name or values");
    REQUIRE(d.dof_per_loc >= 1 && d.dof_per_loc <= 32, "vtu_write: dof_per_loc must be 1..32");
    REQUIRE(d.data_type >= 0 && d.data_type <= 2 && (d.location == 0 || d.location == 1),
            "vtu_write: bad data type or location");
    int nc = 0;
    for (int c = 0; c < d.dof_per_loc; ++c) nc += (d.mask >> c) & 1u;
    REQUIRE(nc >= 1, std::string("vtu_write: empty dof mask: data = ") + d.name);
    REQUIRE(!(nc > 1 && d.data_type == 0),
            std::string("VTUBackend::write_data: Incompatible OutputData::dof_mask and OutputData::DataType: data = ") +
                d.name);
    REQUIRE(nc <= 3, std::string("vtu_write: at most three components per array (the reference pads vectors and "
                                 "tensors to three): data = ") + d.name);
    REQUIRE(d.location == 0 || d.dof_per_loc == 1, "vtu_write: cell data with one value per cell only");
    const int64_t nloc = d.location == 0 ? n_points : n_cells;
    const size_t len = (size_t)(nloc * d.dof_per_loc);
    vals[k].resize(len);
    if (len) {
      if (d.on_device) HIPCHK(hipMemcpy(vals[k].data(), d.data, len * 8, hipMemcpyDeviceToHost));
      else std::memcpy(vals[k].data(), d.data, len * 8);
    }
  }
  const std::string fname = std::string(fname_base) + ".vtu";
  VtuOut out;
  out.f = std::fopen(fname.c_str(), "w");
  REQUIRE(out.f, "vtu_write: cannot open " + fname);
  std::FILE *f = out.f;
  const int64_t n_out = n_cells * nsub;
  std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\">\n");
  std::fprintf(f, "<UnstructuredGrid>\n<Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n", (long long)n_points,
               (long long)n_out);
  std::fprintf(f, "<Points>\n<DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (int64_t i = 0; i < n_points; ++i) {
    put_num(f, points[3 * i]);
    std::fputc(' ', f);
    put_num(f, points[3 * i + 1]);
    std::fputc(' ', f);
    put_num(f, points[3 * i + 2]);
    std::fputc('\n', f);
  }
  std::fprintf(f, "</DataArray>\n</Points>\n<Cells>\n<DataArray type=\"UInt32\" Name=\"connectivity\" format=\"ascii\">\n");
  int sub[8];
  for (int64_t e = 0; e < n_cells; ++e) {
    const int32_t *g = &cells[e * nen];
    for (int s = 0; s < nsub; ++s) {
      if (dim == 2) {
        const int i = s / p, j = s % p;
        sub[0] = i * n1 + j;
        sub[1] = i * n1 + j + 1;
        sub[2] = (i + 1) * n1 + j + 1;
        sub[3] = (i + 1) * n1 + j;
      } else {
        const int i = s / (p * p), j = (s / p) % p, k = s % p;
        const int c = i * n1 * n1 + j * n1 + k;
        sub[0] = c;
        sub[1] = c + 1;
        sub[2] = c + 1 + n1;
        sub[3] = c + n1;
        for (int t = 0; t < 4; ++t) sub[4 + t] = sub[t] + n1 * n1;
      }
      for (int t = 0; t < nv; ++t) std::fprintf(f, t + 1 < nv ? "%d " : "%d\n", g[sub[t]]);
    }
  }
  std::fprintf(f, "</DataArray>\n<DataArray type=\"UInt32\" Name=\"offsets\" format=\"ascii\">\n");
  for (int64_t c = 1; c <= n_out; ++c) std::fprintf(f, "%lld\n", (long long)(c * nv));
  std::fprintf(f, "</DataArray>\n<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
  for (int64_t c = 0; c < n_out; ++c) std::fprintf(f, "%d\n", dim == 2 ? 9 : 12);
  std::fprintf(f, "</DataArray>\n</Cells>\n");
  bool any = false;
  for (int k = 0; k < n_data; ++k) any = any || data[k].location == 0;
  if (any) {
    std::fprintf(f, "<PointData>\n");
    for (int k = 0; k < n_data; ++k)
      if (data[k].location == 0) write_array(f, data[k], vals[k], n_points);
    std::fprintf(f, "</PointData>\n");
  }
  any = false;
  for (int k = 0; k < n_data; ++k) any = any || data[k].location == 1;
  if (any) {
    std::fprintf(f, "<CellData>\n");
    for (int k = 0; k < n_data; ++k) {
      if (data[k].location != 1) continue;
      std::vector<double> rep((size_t)n_out);
      for (int64_t c = 0; c < n_cells; ++c)
        for (int s = 0; s < nsub; ++s) rep[(size_t)(c * nsub + s)] = vals[k][(size_t)c];
      write_array(f, data[k], rep, n_out);
    }
    std::fprintf(f, "</CellData>\n");
  }
  std::fprintf(f, "</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
  REQUIRE(std::ferror(f) == 0, "vtu_write: write error on " + fname);
  return YAFEL_HIP_OK;
  YH_TRY_END(nullptr)
}
