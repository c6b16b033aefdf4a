// fsk_sort.hip - device radix sorts of FP32 keys (hipCUB), used to put 1-D
// points and queries in ascending order so that a block of queries meets
// only the data points whose kernel terms can be nonzero.
//
// Built without the library's device-only "-mno-amdgpu-ieee -fno-honor-nans"
// (hipcc 7.2 crashes on hipCUB's radix sort with them; the Makefile gives
// this file its own flags). The sorts are stable and deterministic; their
// scratch comes from the caller's workspace.
#include <hipcub/hipcub.hpp>

#include "fsk_common.h"

namespace fsk {

size_t sort_f32_temp_bytes(int64_t n, bool pairs) {
  if (n <= 0 || n > INT32_MAX) return 0;
  size_t bytes = 0;
  hipError_t e;
  if (pairs)
    e = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, static_cast<const float *>(nullptr),
                                           static_cast<float *>(nullptr),
                                           static_cast<const int *>(nullptr),
                                           static_cast<int *>(nullptr), int(n), 0, 32,
                                           hipStream_t(0));
  else
    e = hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, static_cast<const float *>(nullptr),
                                          static_cast<float *>(nullptr), int(n), 0, 32,
                                          hipStream_t(0));
  return e == hipSuccess ? bytes : 0;
}

int sort_keys_f32(const char *what, const float *in, float *out, int64_t n, void *tmp,
                  size_t tmp_bytes, hipStream_t s) {
  size_t bytes = tmp_bytes;
  if (hipcub::DeviceRadixSort::SortKeys(tmp, bytes, in, out, int(n), 0, 32, s) != hipSuccess) {
    set_error("%s: radix sort of %lld keys failed", what, (long long)n);
    return FSK_ELAUNCH;
  }
  return launch_status(what);
}

int sort_pairs_f32_i32(const char *what, const float *kin, float *kout, const int *vin,
                       int *vout, int64_t n, void *tmp, size_t tmp_bytes, hipStream_t s) {
  size_t bytes = tmp_bytes;
  if (hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, int(n), 0, 32, s) !=
      hipSuccess) {
    set_error("%s: radix sort of %lld pairs failed", what, (long long)n);
    return FSK_ELAUNCH;
  }
  return launch_status(what);
}

}  // namespace fsk
