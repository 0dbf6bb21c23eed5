// YAFEL's FE layer over the HIP C ABI: Mesh, DoFManager, FESystem,
// CGAssembly<Physics>, DGAssembly<Physics>, DGRK4 and DirichletBC with the
// reference's names.
//
//   Mesh(fname)                           src/mesh/Mesh.cpp:32-36, src/mesh/parse_gmsh.cpp:80-151 (gmsh 2.2)
//   DoFManager(M, CG|DG, p, dof_per_node) src/utils/DoFManager.cpp:11-36 (numbering on the device)
//   FESystem(dofm)                        include/fe_system/FESystem.hpp:18-56
//   CGAssembly<Physics>(fesys, reqs)      include/assembly/CGAssembly.hpp:27-186
//   DGAssembly<Physics>(fesys, P, reqs)   include/assembly/DGAssembly.hpp:32-305
//   DGRK4(dt).step(fesys, P)              include/time_integration/DGRK4.hpp:33-69
//   DirichletBC(dofm, value, comp)        include/boundary_conditions/DirichletBC.hpp:20-89
//   LocalSmoothingGradient<Physics>(fesys) include/assembly/LocalSmoothingGradient.hpp:20-94
//   ZZGradientRecovery<Physics>(fesys)     include/assembly/ZZGradientRecovery.hpp:33-141
//
// A GPU cannot call a host Physics functor, so a Physics type provides
//     static yafel_hip_physics device_descriptor();
// (kind + coefficients) instead of LocalTangent / LocalResidual.  A Physics
// without it does not compile: there is no CPU fallback.
#pragma once

#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "yafel_amd/linalg.hpp"

namespace yafel {

struct coordinate {
  double x[3] = {0, 0, 0};
  double operator()(int i) const { return x[i]; }
  double &operator()(int i) { return x[i]; }
};

// Explicit mesh (geometry nodes + uniform Quad4 / Hex8 cells), the input the
// reference's Mesh(nodes, cells, offsets, types) constructor takes
// (src/mesh/Mesh.cpp:11-29).  Corner order: lexicographic for p = 1,
// gmsh counter-clockwise for p >= 2 (see yafel_hip_mesh_create_cg).
class Mesh {
 public:
  Mesh(int dim, std::vector<double> nodes_xyz, std::vector<int32_t> cells)
      : dim_(dim), nodes_(std::move(nodes_xyz)), cells_(std::move(cells)) {}
  // Mesh(const std::string &fname) (src/mesh/Mesh.cpp:32-36): a gmsh 2.2 file.
  // The volume cells (Hex8 if the file has any, else Quad4) are kept in file
  // order; lines, points and (in 3-D) faces are boundary tags that
  // CGAssembly skips (topoDim != nsd).  Simplex volume cells are not covered
  // by the device elements: std::invalid_argument.
  explicit Mesh(const std::string &fname) {
    yafel_hip_gmsh g = nullptr;
    hipdetail::check(yafel_hip_gmsh_read(fname.c_str(), &g));
    std::shared_ptr<yafel_hip_gmsh_s> hold(g, [](yafel_hip_gmsh_s *p) { yafel_hip_gmsh_destroy(p); });
    int64_t sz[4];
    hipdetail::check(yafel_hip_gmsh_sizes(g, sz));
    std::vector<int32_t> cn((std::size_t)sz[2]), off((std::size_t)sz[1] + 1), ty((std::size_t)sz[1]);
    nodes_.resize((std::size_t)sz[0] * 3);
    hipdetail::check(yafel_hip_gmsh_arrays(g, nodes_.data(), cn.data(), off.data(), ty.data(), nullptr, nullptr));
    bool hex = false, tri = false, tet = false;
    for (int t : ty) {
      hex = hex || t == YAFEL_HIP_CELL_HEX8;
      tri = tri || t == YAFEL_HIP_CELL_TRI3;
      tet = tet || t == YAFEL_HIP_CELL_TET4;
    }
    if (tet || (tri && !hex))
      throw std::invalid_argument("Mesh: simplex volume cells are not covered by the device elements");
    dim_ = hex ? 3 : 2;
    const int want = hex ? YAFEL_HIP_CELL_HEX8 : YAFEL_HIP_CELL_QUAD4, nen = hex ? 8 : 4;
    for (std::size_t c = 0; c < ty.size(); ++c) {
      if (ty[c] != want) continue;
      if (off[c + 1] - off[c] != nen) throw std::invalid_argument("Mesh: cell with a wrong node count");
      cells_.insert(cells_.end(), cn.begin() + off[c], cn.begin() + off[c + 1]);
    }
  }
  // unit square / cube with nx*ny(*nz) cells, corners ordered for polynomial order p
  static Mesh structured(int dim, int nx, int ny, int nz, int p) {
    const std::size_t nn = (std::size_t)(nx + 1) * (ny + 1) * (dim == 3 ? nz + 1 : 1);
    const std::size_t nc = (std::size_t)nx * ny * (dim == 3 ? nz : 1);
    std::vector<double> nodes(3 * nn);
    std::vector<int32_t> cells(nc * (dim == 3 ? 8 : 4));
    hipdetail::check(yafel_hip_structured_mesh(dim, nx, ny, nz, p, nodes.data(), cells.data()));
    return Mesh(dim, std::move(nodes), std::move(cells));
  }
  int dim() const { return dim_; }
  std::size_t nNodes() const { return nodes_.size() / 3; }
  std::size_t nCells() const { return cells_.size() / (dim_ == 3 ? 8 : 4); }
  const std::vector<double> &nodes() const { return nodes_; }
  const std::vector<int32_t> &cells() const { return cells_; }

 private:
  int dim_;
  std::vector<double> nodes_;
  std::vector<int32_t> cells_;
};

class DoFManager {
 public:
  enum class ManagerType { CG, DG };

  DoFManager(const Mesh &M, ManagerType type, int polyOrder, int dof_per_node)
      : dof_per_node(dof_per_node), polyOrder(polyOrder), type_(type) {
    auto ctx = hipdetail::context();
    if (type == ManagerType::DG) {
      // make_dg_dofs (DoFManager.cpp:166-174): element-private nodes, matched faces
      if (dof_per_node != 1 || M.dim() != 2)
        throw std::invalid_argument("DoFManager: the device DG numbering covers 2-D meshes with 1 dof per node");
      yafel_hip_dg g = nullptr;
      hipdetail::check(yafel_hip_dg_create(ctx, polyOrder, (int64_t)M.nNodes(), M.nodes().data(),
                                           (int64_t)M.nCells(), M.cells().data(), &g),
                       ctx);
      dg_ = std::shared_ptr<yafel_hip_dg_s>(g, [](yafel_hip_dg_s *p) { yafel_hip_dg_destroy(p); });
      int64_t info[4];
      hipdetail::check(yafel_hip_dg_info(g, info));
      nen_ = (int)info[3];
      dim_ = 2;
      void *xd = nullptr;
      hipdetail::check(yafel_hip_malloc(ctx, sizeof(double) * 3 * (size_t)info[0], &xd), ctx);
      std::vector<double> xyz((std::size_t)info[0] * 3);
      int rc = yafel_hip_dg_nodes(g, (double *)xd);
      if (rc == YAFEL_HIP_OK) rc = yafel_hip_memcpy_d2h(ctx, xyz.data(), xd, sizeof(double) * xyz.size());
      yafel_hip_free(ctx, xd);
      hipdetail::check(rc, ctx);
      const std::size_t ncell = (std::size_t)info[0] / nen_;
      elements.resize((std::size_t)info[0]);
      for (std::size_t i = 0; i < elements.size(); ++i) elements[i] = (int)i;
      set_nodes(xyz, ncell);
      return;
    }
    yafel_hip_mesh m = nullptr;
    hipdetail::check(yafel_hip_mesh_create_cg(ctx, M.dim(), polyOrder, dof_per_node, (int64_t)M.nNodes(),
                                              M.nodes().data(), (int64_t)M.nCells(), M.cells().data(), &m),
                     ctx);
    m_ = std::shared_ptr<yafel_hip_mesh_s>(m, [](yafel_hip_mesh_s *p) { yafel_hip_mesh_destroy(p); });
    int64_t info[8];
    hipdetail::check(yafel_hip_mesh_info(m, info));
    nen_ = (int)info[2];
    dim_ = (int)info[4];
    elements.resize((std::size_t)info[1] * nen_);
    std::vector<double> xyz((std::size_t)info[0] * 3);
    hipdetail::check(yafel_hip_mesh_download(m, elements.data(), xyz.data()));
    set_nodes(xyz, (std::size_t)info[1]);
  }

  int nCells() const { return (int)(element_offsets.size() - 1); }
  int nsd() const { return dim_; }
  ManagerType managerType() const { return type_; }
  // DoFManager::getGlobalDofs (DoFManager.cpp:39-52): dof = node * dpn + comp
  void getGlobalDofs(int elnum, std::vector<int> &c) const {
    c.clear();
    for (int i = element_offsets[elnum]; i < element_offsets[elnum + 1]; ++i)
      for (int d = 0; d < dof_per_node; ++d) c.push_back(elements[i] * dof_per_node + d);
  }
  void getGlobalNodes(int elnum, std::vector<int> &c) const {
    c.assign(elements.begin() + element_offsets[elnum], elements.begin() + element_offsets[elnum + 1]);
  }
  yafel_hip_mesh handle() const { return m_.get(); }
  yafel_hip_dg dg_handle() const { return dg_.get(); }

  int dof_per_node;
  int polyOrder;
  std::vector<int> elements;
  std::vector<int> element_offsets;
  std::vector<coordinate> dof_nodes;

 private:
  void set_nodes(const std::vector<double> &xyz, std::size_t ncell) {
    dof_nodes.resize(xyz.size() / 3);
    for (std::size_t i = 0; i < dof_nodes.size(); ++i)
      for (int d = 0; d < 3; ++d) dof_nodes[i].x[d] = xyz[3 * i + d];
    element_offsets.resize(ncell + 1);
    for (std::size_t e = 0; e < element_offsets.size(); ++e) element_offsets[e] = (int)(e * nen_);
  }
  ManagerType type_;
  std::shared_ptr<yafel_hip_mesh_s> m_;
  std::shared_ptr<yafel_hip_dg_s> dg_;
  int nen_ = 0, dim_ = 0;
};

enum class AssemblyRequirement : int { Residual, Tangent, DtMass, LumpedDtMassInverse, DtDtMass };

// Owns the device tangent (the mesh's sparsity pattern), residual and the host
// solution.  format: YAFEL_HIP_CSR or YAFEL_HIP_BCSR3 (3 dofs per node).
class FESystem {
 public:
  explicit FESystem(DoFManager &dofm, int format = YAFEL_HIP_CSR) : dofm_(dofm) {
    auto ctx = hipdetail::context();
    if (dofm.managerType() == DoFManager::ManagerType::CG) {
      // DG systems are explicit: no tangent
      yafel_hip_matrix A = nullptr;
      hipdetail::check(yafel_hip_mesh_pattern(dofm.handle(), format, YAFEL_HIP_FP64, &A), ctx);
      A_ = hipdetail::MatrixPtr(A, hipdetail::MatrixDeleter());
    }
    n_ = dofm.dof_nodes.size() * dofm.dof_per_node;
    void *r = nullptr;
    hipdetail::check(yafel_hip_malloc(ctx, sizeof(double) * n_, &r), ctx);
    R_ = std::shared_ptr<double>((double *)r, [ctx](double *p) { yafel_hip_free(ctx, p); });
    std::vector<double> z(n_, 0.0);
    hipdetail::check(yafel_hip_memcpy_h2d(ctx, r, z.data(), sizeof(double) * n_), ctx);
    solution_ = Vector<double>(n_, 0.0);
  }

  yafel_hip_matrix tangent_handle() const { return A_.get(); }
  double *residual_device() const { return R_.get(); }
  std::size_t ndofs() const { return n_; }
  DoFManager &getDoFManager() { return dofm_; }
  double &currentTime() { return time_; }

  // host copies of the device tangent / residual
  sparse_csr<double> getGlobalTangent() const { return sparse_csr<double>(A_.get()); }
  sparse_bcsr<3, double> getGlobalTangentBlocks() const { return sparse_bcsr<3, double>(A_.get()); }
  Vector<double> getGlobalResidual() const {
    Vector<double> r(n_, 0.0);
    auto ctx = hipdetail::context();
    hipdetail::check(yafel_hip_memcpy_d2h(ctx, r.data(), R_.get(), sizeof(double) * n_), ctx);
    return r;
  }
  Vector<double> &getSolution() { return solution_; }
  // RowMajor n_dofs x nsd (n_nodes x nsd after ZZ recovery), as the
  // reference's solution_gradient (include/fe_system/FESystem.hpp:37, 53)
  std::vector<double> &getSolutionGradient() { return gradient_; }

  // solve tangent * U = residual on the device (b stays resident)
  yafel_hip_solve_info solve(int method = YAFEL_HIP_PCG_JACOBI, int precision = YAFEL_HIP_FP64,
                             double tol = 1e-8) {
    auto ctx = hipdetail::context();
    yafel_hip_solve_opts o;
    yafel_hip_solve_opts_default(&o);
    o.method = method;
    o.precision = precision;
    o.stop = YAFEL_HIP_STOP_RELRES;
    o.tol = tol;
    void *x = nullptr;
    hipdetail::check(yafel_hip_malloc(ctx, sizeof(double) * n_, &x), ctx);
    yafel_hip_solve_info info{};
    int rc = yafel_hip_memcpy_h2d(ctx, x, solution_.data(), sizeof(double) * n_);
    if (rc == YAFEL_HIP_OK) rc = yafel_hip_solve(A_.get(), &o, R_.get(), (double *)x, &info);
    if (rc == YAFEL_HIP_OK) rc = yafel_hip_memcpy_d2h(ctx, solution_.data(), x, sizeof(double) * n_);
    yafel_hip_free(ctx, x);
    hipdetail::check(rc, ctx);
    return info;
  }

 private:
  DoFManager &dofm_;
  hipdetail::MatrixPtr A_;
  std::shared_ptr<double> R_;
  std::size_t n_ = 0;
  Vector<double> solution_;
  std::vector<double> gradient_;
  double time_ = 0;
};

namespace hipdetail {
// the host solution up, the recovered gradient (rows x nsd) back
template <class Fn>
void recover_gradient(FESystem &fes, std::size_t rows, int nsd, Fn &&fn) {
  auto ctx = context();
  const std::size_t n = fes.ndofs();
  void *u = nullptr, *g = nullptr;
  check(yafel_hip_malloc(ctx, sizeof(double) * n, &u), ctx);
  int rc = yafel_hip_malloc(ctx, sizeof(double) * rows * nsd + 8, &g);
  if (rc == YAFEL_HIP_OK) rc = yafel_hip_memcpy_h2d(ctx, u, fes.getSolution().data(), sizeof(double) * n);
  if (rc == YAFEL_HIP_OK) rc = fn((const double *)u, (double *)g);
  auto &out = fes.getSolutionGradient();
  out.assign(rows * nsd, 0.0);
  if (rc == YAFEL_HIP_OK) rc = yafel_hip_memcpy_d2h(ctx, out.data(), g, sizeof(double) * rows * nsd);
  if (g) yafel_hip_free(ctx, g);
  yafel_hip_free(ctx, u);
  check(rc, ctx);
}
}  // namespace hipdetail

// LocalSmoothingGradient<Physics>(FESystem&) (include/assembly/
// LocalSmoothingGradient.hpp:20-94): the volume-weighted nodal average of the
// solution gradient, per dof, into getSolutionGradient() (n_dofs x nsd).
template <typename Physics>
void LocalSmoothingGradient(FESystem &fes) {
  constexpr int NSD = Physics::nsd();
  yafel_hip_mesh m = fes.getDoFManager().handle();
  hipdetail::recover_gradient(fes, fes.ndofs(), NSD, [m](const double *u, double *g) {
    return yafel_hip_gradient_local_smoothing(m, u, g);
  });
}

// ZZGradientRecovery<Physics>(FESystem&) (include/assembly/ZZGradientRecovery.hpp:
// 33-141): p = 1 scalar fields; the nodal least-squares patch fit of the
// quadrature-point gradients into getSolutionGradient() (n_nodes x nsd).
template <typename Physics>
void ZZGradientRecovery(FESystem &fes) {
  constexpr int NSD = Physics::nsd();
  yafel_hip_mesh m = fes.getDoFManager().handle();
  hipdetail::recover_gradient(fes, fes.getDoFManager().dof_nodes.size(), NSD, [m](const double *u, double *g) {
    return yafel_hip_gradient_zz(m, u, g);
  });
}

// CGAssembly<Physics> (include/assembly/CGAssembly.hpp:27-186): tangent values
// into the fixed pattern (overwritten), residual ACCUMULATED.
template <typename Physics>
void CGAssembly(FESystem &fes, std::vector<AssemblyRequirement> requirements = {AssemblyRequirement::Residual,
                                                                                AssemblyRequirement::Tangent}) {
  int req = 0;
  for (auto r : requirements) {
    if (r == AssemblyRequirement::Residual) req |= YAFEL_HIP_ASSEMBLE_RESIDUAL;
    else if (r == AssemblyRequirement::Tangent) req |= YAFEL_HIP_ASSEMBLE_TANGENT;
    else throw std::invalid_argument("CGAssembly: mass requirements are not part of the device path");
  }
  const yafel_hip_physics ph = Physics::device_descriptor();
  hipdetail::check(yafel_hip_assemble(fes.getDoFManager().handle(), &ph, req, fes.tangent_handle(),
                                      fes.residual_device()),
                   hipdetail::context());
}

// DGAssembly<Physics> (include/assembly/DGAssembly.hpp:32-305) with
// {Residual, DtMass}: the residual of the host solution into the device
// residual (overwritten, as GlobalResidual *= 0 does); the element mass
// inverses live with the DoFManager's device state, built once.
template <typename Physics>
void DGAssembly(FESystem &fes, Physics &, std::vector<AssemblyRequirement> requirements =
                                               {AssemblyRequirement::Residual}) {
  for (auto r : requirements)
    if (r == AssemblyRequirement::Tangent || r == AssemblyRequirement::DtDtMass)
      throw std::invalid_argument("DGAssembly: the device path assembles the explicit residual");
  auto ctx = hipdetail::context();
  const yafel_hip_physics ph = Physics::device_descriptor();
  const std::size_t n = fes.ndofs();
  void *u = nullptr;
  hipdetail::check(yafel_hip_malloc(ctx, sizeof(double) * n, &u), ctx);
  int rc = yafel_hip_memcpy_h2d(ctx, u, fes.getSolution().data(), sizeof(double) * n);
  if (rc == YAFEL_HIP_OK)
    rc = yafel_hip_dg_residual(fes.getDoFManager().dg_handle(), &ph, fes.currentTime(), (const double *)u,
                               fes.residual_device(), nullptr);
  if (rc == YAFEL_HIP_OK) rc = yafel_hip_ctx_sync(ctx);
  yafel_hip_free(ctx, u);
  hipdetail::check(rc, ctx);
}

// DGRK4 (include/time_integration/DGRK4.hpp:22-69): classic RK4 on
// dU/dt = DGAssembly residual.  step() advances the host solution by dt;
// run() keeps the solution on the device for nsteps replays of the captured
// step graph.
class DGRK4 {
 public:
  DGRK4() = delete;
  explicit DGRK4(double dt) : dt(dt) {}

  template <typename Physics>
  void step(FESystem &fes, Physics &P) {
    run(fes, P, 1);
  }

  template <typename Physics>
  void run(FESystem &fes, Physics &, int64_t nsteps) {
    auto ctx = hipdetail::context();
    const yafel_hip_physics ph = Physics::device_descriptor();
    const std::size_t n = fes.ndofs();
    void *u = nullptr;
    hipdetail::check(yafel_hip_malloc(ctx, sizeof(double) * n, &u), ctx);
    double t = fes.currentTime();
    int rc = yafel_hip_memcpy_h2d(ctx, u, fes.getSolution().data(), sizeof(double) * n);
    if (rc == YAFEL_HIP_OK) rc = yafel_hip_dg_rk4(fes.getDoFManager().dg_handle(), &ph, dt, nsteps, &t, (double *)u);
    if (rc == YAFEL_HIP_OK) rc = yafel_hip_memcpy_d2h(ctx, fes.getSolution().data(), u, sizeof(double) * n);
    yafel_hip_free(ctx, u);
    hipdetail::check(rc, ctx);
    fes.currentTime() = t;
  }

 private:
  double dt;
};

// DirichletBC (DirichletBC.hpp:20-89, DirichletBC.cpp:27-56)
class DirichletBC {
 public:
  DirichletBC(const DoFManager &dofm, double bcval, int comp = 0)
      : dofm_(dofm), comp_(comp), value_([bcval](const coordinate &, double) { return bcval; }) {}
  DirichletBC(const DoFManager &dofm, std::function<double(const coordinate &, double)> f, int comp = 0)
      : dofm_(dofm), comp_(comp), value_(std::move(f)) {}

  template <typename Lambda>
  void selectByFunction(Lambda &&func) {
    for (std::size_t i = 0; i < dofm_.dof_nodes.size(); ++i)
      if (func(dofm_.dof_nodes[i])) nodes_.push_back((int)i);
  }
  std::vector<int> &getNodes() { return nodes_; }

  void apply(FESystem &fes, double time = 0.0) {
    const std::size_t n = fes.ndofs();
    std::vector<uint8_t> mask(n, 0);
    std::vector<double> g(n, 0.0);
    for (int node : nodes_) {
      const std::size_t d = (std::size_t)node * dofm_.dof_per_node + comp_;
      mask[d] = 1;
      g[d] = value_(dofm_.dof_nodes[node], time);
    }
    auto ctx = hipdetail::context();
    void *dm = nullptr, *dg = nullptr;
    hipdetail::check(yafel_hip_malloc(ctx, n, &dm), ctx);
    hipdetail::check(yafel_hip_malloc(ctx, sizeof(double) * n, &dg), ctx);
    int rc = yafel_hip_memcpy_h2d(ctx, dm, mask.data(), n);
    if (rc == YAFEL_HIP_OK) rc = yafel_hip_memcpy_h2d(ctx, dg, g.data(), sizeof(double) * n);
    if (rc == YAFEL_HIP_OK)
      rc = yafel_hip_dirichlet_apply(fes.tangent_handle(), (const uint8_t *)dm, (const double *)dg,
                                     fes.residual_device());
    if (rc == YAFEL_HIP_OK) rc = yafel_hip_ctx_sync(ctx);
    yafel_hip_free(ctx, dm);
    yafel_hip_free(ctx, dg);
    hipdetail::check(rc, ctx);
  }

 private:
  const DoFManager &dofm_;
  int comp_;
  std::function<double(const coordinate &, double)> value_;
  std::vector<int> nodes_;
};

}  // namespace yafel
