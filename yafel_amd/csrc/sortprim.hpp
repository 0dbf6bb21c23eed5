// Device-wide sort/scan helpers (sort.hip).  All calls are stream-ordered and
// return after the stream has drained.
#pragma once
#include "internal.hpp"

namespace yh {
// stable LSD radix sort of (key, value) pairs on the low end_bit key bits, in place
void sort_pairs_u64_u32(uint64_t *keys, uint32_t *vals, int64_t n, int end_bit, hipStream_t s);
void sort_pairs_u64_u64(uint64_t *keys, uint64_t *vals, int64_t n, int end_bit, hipStream_t s);
void exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s);
void inclusive_scan_i32(const int32_t *in, int32_t *out, int64_t n, hipStream_t s);

inline int bits_for(uint64_t max_value) {
  int b = 1;
  while (b < 64 && (max_value >> b) != 0) ++b;
  return b;
}

// matrix_from_device (sparse.hip): takes ownership of the device arrays
yafel_hip_matrix_s *matrix_from_device(yafel_hip_ctx_s *ctx, int format, int64_t brows, int64_t bcols,
                                       int64_t nblk, int64_t ptr_len, DBuf &ptr, DBuf &col,
                                       DBuf &val64, int precisions);
}  // namespace yh
