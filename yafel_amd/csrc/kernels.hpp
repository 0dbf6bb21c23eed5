// Device building blocks shared by the SpMV and Krylov kernels.
#pragma once
#include "internal.hpp"

namespace yh {
namespace dev {

// One CSR row with L lanes: lane l sums entries l, l+L, ... of the row, then
// the L partial sums are combined with xor-shuffles.  Matrix values and
// column indices are streamed once, so they are loaded non-temporally; the
// gathered vector stays in cache.  xf(c) returns x[c].
// T is the stored value type, TV the vector / accumulation type.
template <int L, class T, class TV, class XF>
__device__ __forceinline__ TV csr_row(const int *__restrict__ ptr, const int *__restrict__ col,
                                      const T *__restrict__ val, int64_t r, int lane, XF xf) {
  const int b = ptr[r], e = ptr[r + 1];
  TV acc = 0, acc2 = 0;
  int k = b + lane;
  for (; k + L < e; k += 2 * L) {
    const int c = __builtin_nontemporal_load(&col[k]);
    const int c2 = __builtin_nontemporal_load(&col[k + L]);
    const T v = __builtin_nontemporal_load(&val[k]);
    const T v2 = __builtin_nontemporal_load(&val[k + L]);
    acc = fma((TV)v, xf(c), acc);
    acc2 = fma((TV)v2, xf(c2), acc2);
  }
  if (k < e) acc = fma((TV)__builtin_nontemporal_load(&val[k]), xf(__builtin_nontemporal_load(&col[k])), acc);
  return group_sum<L>(acc + acc2);
}

// One 3x3-block row with G lanes, lane per block; values are entry-major
// (val[e * nb + k]).  a[0..2] receive the three row sums (valid in every lane
// of the group).  xf(i) returns x[i] for a scalar index.
template <int G, class T, class TV, class XF>
__device__ __forceinline__ void bcsr_row(const int *__restrict__ ptr, const int *__restrict__ col,
                                         const T *__restrict__ val, int64_t nb, int64_t br,
                                         int lane, TV a[3], XF xf) {
  const int b = ptr[br], e = ptr[br + 1];
  TV a0 = 0, a1 = 0, a2 = 0;
  for (int k = b + lane; k < e; k += G) {
    const int64_t c = 3 * (int64_t)__builtin_nontemporal_load(&col[k]);
    const TV x0 = xf(c), x1 = xf(c + 1), x2 = xf(c + 2);
    a0 = fma((TV)__builtin_nontemporal_load(&val[0 * nb + k]), x0, a0);
    a0 = fma((TV)__builtin_nontemporal_load(&val[1 * nb + k]), x1, a0);
    a0 = fma((TV)__builtin_nontemporal_load(&val[2 * nb + k]), x2, a0);
    a1 = fma((TV)__builtin_nontemporal_load(&val[3 * nb + k]), x0, a1);
    a1 = fma((TV)__builtin_nontemporal_load(&val[4 * nb + k]), x1, a1);
    a1 = fma((TV)__builtin_nontemporal_load(&val[5 * nb + k]), x2, a1);
    a2 = fma((TV)__builtin_nontemporal_load(&val[6 * nb + k]), x0, a2);
    a2 = fma((TV)__builtin_nontemporal_load(&val[7 * nb + k]), x1, a2);
    a2 = fma((TV)__builtin_nontemporal_load(&val[8 * nb + k]), x2, a2);
  }
  a[0] = group_sum<G>(a0);
  a[1] = group_sum<G>(a1);
  a[2] = group_sum<G>(a2);
}

}  // namespace dev
}  // namespace yh
