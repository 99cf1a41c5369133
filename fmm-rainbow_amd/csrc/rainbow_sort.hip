// Radix sort of (u32 key, u32 value) pairs, used to keep bodies in Morton
// order of their finest cell.  rocPRIM's device radix sort (header-only).
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "rainbow_launch.h"

namespace rainbow {

hipError_t sort_pairs(void* temp, size_t& temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                      const uint32_t* vals_in, uint32_t* vals_out, uint32_t n, uint32_t end_bit, hipStream_t stream) {
  // Onesweep radix sort at every size above one block: the default switches
  // to a merge sort up to 2^20 items, which is several times slower here.
  using config = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                            rocprim::default_config, 0>;
  return rocprim::radix_sort_pairs<config>(temp, temp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0u, end_bit,
                                           stream);
}

}  // namespace rainbow
