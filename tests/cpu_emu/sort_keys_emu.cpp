// CPU stand-in for csrc/sort_keys.hip under the emulation (device memory is
// host memory there): the same stable sort of (key, value) pairs on the low
// key_bits bits, done with std::stable_sort.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <vector>

namespace rpd {

size_t sort_pairs_u32_temp_bytes(int) { return 16; }

hipError_t sort_pairs_u32(const uint32_t* keys_in, uint32_t* keys_out, const int* vals_in, int* vals_out, int n,
                          int key_bits, void*, size_t, hipStream_t) {
    const uint32_t mask = key_bits >= 32 ? 0xffffffffu : ((1u << key_bits) - 1u);
    std::vector<int> order(size_t(std::max(n, 0)));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return (keys_in[a] & mask) < (keys_in[b] & mask); });
    std::vector<uint32_t> k(order.size());
    std::vector<int> v(order.size());
    for (size_t i = 0; i < order.size(); ++i) { k[i] = keys_in[order[i]]; v[i] = vals_in[order[i]]; }
    std::copy(k.begin(), k.end(), keys_out);
    std::copy(v.begin(), v.end(), vals_out);
    return hipSuccess;
}

}  // namespace rpd
