// Device-wide primitives (stable LSD radix sort, scans) used to build
// sparsity patterns, assembly gather plans and DoF numberings on the device.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "internal.hpp"
#include "sortprim.hpp"

namespace yh {

// The temporaries are released when a wrapper returns.  Stream-ordered ones
// (a setup thread's allocation stream, the sort's stream) are freed in
// stream order; plain ones need the sort to have finished first.
static void finish_temps(const DBuf &t, hipStream_t s) {
  if (t.s != s) HIPCHK(hipStreamSynchronize(s));
}

void sort_pairs_u64_u32(uint64_t *keys, uint32_t *vals, int64_t n, int end_bit, hipStream_t s) {
  if (n <= 1) return;
  DBuf kt, vt, tmp;
  kt.alloc(n * 8);
  vt.alloc(n * 4);
  size_t tb = 0;
  HIPCHK(rocprim::radix_sort_pairs(nullptr, tb, keys, kt.as<uint64_t>(), vals, vt.as<uint32_t>(),
                                   (size_t)n, 0u, (unsigned)end_bit, s));
  tmp.alloc(tb + 16);
  HIPCHK(rocprim::radix_sort_pairs(tmp.p, tb, keys, kt.as<uint64_t>(), vals, vt.as<uint32_t>(),
                                   (size_t)n, 0u, (unsigned)end_bit, s));
  HIPCHK(hipMemcpyAsync(keys, kt.p, n * 8, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(vals, vt.p, n * 4, hipMemcpyDeviceToDevice, s));
  finish_temps(kt, s);
}

void sort_pairs_u64_u64(uint64_t *keys, uint64_t *vals, int64_t n, int end_bit, hipStream_t s) {
  if (n <= 1) return;
  DBuf kt, vt, tmp;
  kt.alloc(n * 8);
  vt.alloc(n * 8);
  size_t tb = 0;
  HIPCHK(rocprim::radix_sort_pairs(nullptr, tb, keys, kt.as<uint64_t>(), vals, vt.as<uint64_t>(),
                                   (size_t)n, 0u, (unsigned)end_bit, s));
  tmp.alloc(tb + 16);
  HIPCHK(rocprim::radix_sort_pairs(tmp.p, tb, keys, kt.as<uint64_t>(), vals, vt.as<uint64_t>(),
                                   (size_t)n, 0u, (unsigned)end_bit, s));
  HIPCHK(hipMemcpyAsync(keys, kt.p, n * 8, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(vals, vt.p, n * 8, hipMemcpyDeviceToDevice, s));
  finish_temps(kt, s);
}

void exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  DBuf tmp;
  size_t tb = 0;
  HIPCHK(rocprim::exclusive_scan(nullptr, tb, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), s));
  tmp.alloc(tb + 16);
  HIPCHK(rocprim::exclusive_scan(tmp.p, tb, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), s));
  finish_temps(tmp, s);
}

void inclusive_scan_i32(const int32_t *in, int32_t *out, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  DBuf tmp;
  size_t tb = 0;
  HIPCHK(rocprim::inclusive_scan(nullptr, tb, in, out, (size_t)n, rocprim::plus<int32_t>(), s));
  tmp.alloc(tb + 16);
  HIPCHK(rocprim::inclusive_scan(tmp.p, tb, in, out, (size_t)n, rocprim::plus<int32_t>(), s));
  finish_temps(tmp, s);
}

}  // namespace yh
