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

// Stable LSD radix sorts on double buffers: the sorted data end in either buffer, and the
// functions return true when they ended in the second (the caller swaps its buffers), which
// saves the copy back.
template <class K, class V>
static bool sort_pairs(Ctx &c, K *keys, K *keys_tmp, V *vals, V *vals_tmp, long n, int end_bit) {
    if (n <= 1) return false;
    rocprim::double_buffer<K> kb(keys, keys_tmp);
    rocprim::double_buffer<V> vb(vals, vals_tmp);
    size_t bytes = 0;
    HBP_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, kb, vb, (size_t)n, 0u, (unsigned)end_bit, c.stream));
    void *tmp = c.scratch_bytes(bytes);
    HBP_CHECK(rocprim::radix_sort_pairs(tmp, bytes, kb, vb, (size_t)n, 0u, (unsigned)end_bit, c.stream));
    HBP_REQUIRE((kb.current() == keys_tmp) == (vb.current() == vals_tmp), "sort buffers out of step");
    return kb.current() == keys_tmp;
}

bool sort_pairs_u64(Ctx &c, unsigned long long *keys, unsigned long long *keys_tmp, unsigned int *vals,
                    unsigned int *vals_tmp, long n, int end_bit) {
    return sort_pairs(c, keys, keys_tmp, vals, vals_tmp, n, end_bit);
}

bool sort_pairs_u32(Ctx &c, unsigned int *keys, unsigned int *keys_tmp, int *vals, int *vals_tmp, long n,
                    int end_bit) {
    return sort_pairs(c, keys, keys_tmp, vals, vals_tmp, n, end_bit);
}

bool sort_keys_u64(Ctx &c, unsigned long long *keys, unsigned long long *keys_tmp, long n, int end_bit) {
    if (n <= 1) return false;
    rocprim::double_buffer<unsigned long long> kb(keys, keys_tmp);
    size_t bytes = 0;
    HBP_CHECK(rocprim::radix_sort_keys(nullptr, bytes, kb, (size_t)n, 0u, (unsigned)end_bit, c.stream));
    void *tmp = c.scratch_bytes(bytes);
    HBP_CHECK(rocprim::radix_sort_keys(tmp, bytes, kb, (size_t)n, 0u, (unsigned)end_bit, c.stream));
    return kb.current() == keys_tmp;
}

}  // namespace hbp
