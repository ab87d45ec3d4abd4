// rocprim_yardstick.hip - rocPRIM's device radix sort behind a C ABI, as a
// yardstick for the pipeline's own key/value sort (bench.py times both on
// the same Morton keys). Not part of the build pipeline: libnxb.so does not
// link it.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <stddef.h>
#include <stdint.h>

// Stable sort of (key, value) pairs on key bits [begin_bit, end_bit) with
// rocprim::radix_sort_pairs. With temp == NULL only *temp_bytes is set.
// Returns the hipError_t as an int.
extern "C" int nxb_ys_sort_pairs_u64(const uint64_t* keys_in, uint64_t* keys_out, const uint32_t* vals_in,
                                     uint32_t* vals_out, uint32_t n, uint32_t begin_bit, uint32_t end_bit,
                                     void* temp, size_t* temp_bytes, hipStream_t stream) {
    return (int)rocprim::radix_sort_pairs(temp, *temp_bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n,
                                          begin_bit, end_bit, stream);
}

extern "C" int nxb_ys_sort_pairs_u32(const uint32_t* keys_in, uint32_t* keys_out, const uint32_t* vals_in,
                                     uint32_t* vals_out, uint32_t n, uint32_t begin_bit, uint32_t end_bit,
                                     void* temp, size_t* temp_bytes, hipStream_t stream) {
    return (int)rocprim::radix_sort_pairs(temp, *temp_bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n,
                                          begin_bit, end_bit, stream);
}
