// Device mesh (DoFManager output), element tables and assembly gather plans.
#pragma once
#include <memory>

#include "internal.hpp"

// Deterministic gather plan for one matrix format: the node-pair runs of the
// sorted element pair list.  Run j (a stored node-pair (A,B)) owns the
// contributions contrib[cptr[j] .. cptr[j+1]) -- element-pair indices
// t = e*nen^2 + a*nen + b in increasing t (element order).
struct YhPlan {
  int format = 0;
  int64_t nruns = 0;      // stored node pairs
  int64_t nrow_nodes = 0; // node rows of the pattern (max row + 1)
  yh::DBuf cptr;          // int32 [nruns+1]
  yh::DBuf contrib;       // uint32 [nel*nen^2]
  yh::DBuf run_row;       // int32 [nruns] node row of each run
  yh::DBuf run_col;       // int32 [nruns] node col of each run
  yh::DBuf node_ptr;      // int32 [n_nodes+1] node pattern row pointer
  yh::DBuf scal_pos;      // CSR with dpn > 1: first scalar position of each run's (c=0,c'=0) entry
};

struct yafel_hip_mesh_s {
  yafel_hip_ctx_s *ctx = nullptr;
  int dim = 2, p = 1, dpn = 1, nen = 4, nqp = 4;
  int64_t n_nodes = 0, nel = 0;
  yh::DBuf elements;  // int32 [nel*nen]
  yh::DBuf nodes;     // double [n_nodes*3]
  yh::DBuf tabN, tabdN, tabw;  // element tables: N[q][a], dN[q][a][d], w[q]
  std::vector<double> hN, hdN, hw;
  // residual plan: node incidences sorted by node
  yh::DBuf inc_ptr;   // int32 [n_nodes+1]
  yh::DBuf inc;       // uint32 [nel*nen] = e*nen + a
  bool inc_ok = false;
  std::unique_ptr<YhPlan> plan_csr, plan_bcsr;
  yh::DBuf stageK, stageR;  // element matrix / vector staging
  yh::DBuf stageK32;        // FP32 element matrix staging (matrix-core path)
  yh::DBuf stageG;          // gradient recovery staging (per element values)
  ~yafel_hip_mesh_s() { yh::ctx_release(ctx); }
};

namespace yh {
// element tables on the host: GLL local nodes, Gauss points and Lagrange basis
void element_tables(int dim, int p, std::vector<double> &N, std::vector<double> &dN,
                    std::vector<double> &w, std::vector<double> &xl);
void mesh_build_incidence(yafel_hip_mesh_s *m);
}  // namespace yh
