// Stable device radix sort of (32-bit key, 32-bit value) pairs, through
// hipCUB. Kept in a file of its own so that the rest of the engine depends on
// this one narrow function only.
#include "kernels.h"

#include <hipcub/device/device_radix_sort.hpp>

namespace rpd {

size_t sort_pairs_u32_temp_bytes(int n) {
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, static_cast<const uint32_t*>(nullptr),
                                           static_cast<uint32_t*>(nullptr), static_cast<const int*>(nullptr),
                                           static_cast<int*>(nullptr), n, 0, 32) != hipSuccess)
        return 0;
    return bytes;
}

hipError_t sort_pairs_u32(const uint32_t* keys_in, uint32_t* keys_out, const int* vals_in, int* vals_out, int n,
                          int key_bits, void* temp, size_t temp_bytes, hipStream_t st) {
    return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0, key_bits,
                                              st);
}

}  // namespace rpd
