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
  int32_t offset = 0;
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
      cell_off_file[e] = offset;
      offset += (int32_t)cell_conn[e].size();
    }
  }
  // cells in id order; each id's nodes at its file-order offset (the reference
  // stores cellOffsets_[id-1] = running offset and appends nodes in file order)
  g->cell_nodes.assign((size_t)offset, 0);
  g->cell_offsets.assign((size_t)nelems + 1, offset);
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
