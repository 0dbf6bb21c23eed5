// Multi-RHS Krylov solves (k right-hand sides advanced together).
#include "internal.hpp"

using namespace yh;

extern "C" int yafel_hip_solve_multi(yafel_hip_matrix A, const yafel_hip_solve_opts *opts, int k,
                                     const double *B_dev, double *X_dev, yafel_hip_solve_info *info) {
  YH_TRY_BEGIN
  (void)opts; (void)k; (void)B_dev; (void)X_dev; (void)info;
  fail(YAFEL_HIP_EINVAL, "solve_multi: not available in this build");
  YH_TRY_END(A ? A->ctx : nullptr)
}
