// Device-wide scan and radix sort used by the pyramid and the BP prologue.
// Both run once per level, off the BP sweep loop; they are rocPRIM's device algorithms (stable
// LSD radix sort, look-back scan), whose results are fully determined by their inputs.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "hbp_internal.h"

namespace hbp {

void scan_exclusive_int(Ctx &c, const int *d_in, int *d_out, long n, int *d_total) {
    (void)d_total;
    size_t bytes = 0;
    HBP_CHECK(rocprim::exclusive_scan(nullptr, bytes, d_in, d_out, 0, (size_t)n, rocprim::plus<int>(), c.stream));
    void *tmp = c.scratch_bytes(bytes);
    HBP_CHECK(rocprim::exclusive_scan(tmp, bytes, d_in, d_out, 0, (size_t)n, rocprim::plus<int>(), c.stream));
}

template <class K, class V>
static void sort_pairs(Ctx &c, K *keys, K *keys_tmp, V *vals, V *vals_tmp, long n, int end_bit) {
    if (n <= 1) return;
    size_t bytes = 0;
    HBP_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, keys, keys_tmp, vals, vals_tmp, (size_t)n, 0u,
                                        (unsigned)end_bit, c.stream));
    void *tmp = c.scratch_bytes(bytes);
    HBP_CHECK(rocprim::radix_sort_pairs(tmp, bytes, keys, keys_tmp, vals, vals_tmp, (size_t)n, 0u,
                                        (unsigned)end_bit, c.stream));
    HBP_CHECK(hipMemcpyAsync(keys, keys_tmp, sizeof(K) * (size_t)n, hipMemcpyDeviceToDevice, c.stream));
    HBP_CHECK(hipMemcpyAsync(vals, vals_tmp, sizeof(V) * (size_t)n, hipMemcpyDeviceToDevice, c.stream));
}

void sort_pairs_u64(Ctx &c, unsigned long long *keys, unsigned long long *keys_tmp, unsigned int *vals,
                    unsigned int *vals_tmp, long n, int end_bit) {
    sort_pairs(c, keys, keys_tmp, vals, vals_tmp, n, end_bit);
}

void sort_pairs_u32(Ctx &c, unsigned int *keys, unsigned int *keys_tmp, int *vals, int *vals_tmp, long n,
                    int end_bit) {
    sort_pairs(c, keys, keys_tmp, vals, vals_tmp, n, end_bit);
}

// keys only; the sorted keys end in keys
void sort_keys_u64(Ctx &c, unsigned long long *keys, unsigned long long *keys_tmp, long n, int end_bit) {
    if (n <= 1) return;
    size_t bytes = 0;
    HBP_CHECK(rocprim::radix_sort_keys(nullptr, bytes, keys, keys_tmp, (size_t)n, 0u, (unsigned)end_bit, c.stream));
    void *tmp = c.scratch_bytes(bytes);
    HBP_CHECK(rocprim::radix_sort_keys(tmp, bytes, keys, keys_tmp, (size_t)n, 0u, (unsigned)end_bit, c.stream));
    HBP_CHECK(hipMemcpyAsync(keys, keys_tmp, sizeof(unsigned long long) * (size_t)n, hipMemcpyDeviceToDevice,
                             c.stream));
}

}  // namespace hbp
