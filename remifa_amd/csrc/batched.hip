// Batched no-pivot LU of many small matrices (sizes up to 512).
#include "common.hpp"

extern "C" int remifa_lu_nopiv_batched(remifa_handle_t h, int ctype, int64_t count, const int* d_n,
                                       const int* d_ld, float* const* d_ptrs, int* d_info,
                                       int max_n) {
  if (!h) return -1;
  remifa::set_last_error("remifa_lu_nopiv_batched: not built yet");
  return REMIFA_ERR_UNSUPPORTED;
}

extern "C" int remifa_getrs_nopiv_batched(remifa_handle_t h, int64_t count, const int* d_n,
                                          const int* d_ld, float* const* d_ptrs,
                                          float* const* d_b, int max_n) {
  if (!h) return -1;
  remifa::set_last_error("remifa_getrs_nopiv_batched: not built yet");
  return REMIFA_ERR_UNSUPPORTED;
}
