// Device-wide exclusive scan and stable LSD radix sort, used once per level by the pyramid
// (edge sort of reduce_boundary_dist, segment.cu:1058-1064; the compaction scan of
// compact_clus_label, CRTrees.cu:395) and by the BP prologue (setup_D4 / sort_clus,
// hbp_isp.cu:888, 993; the CSR scans, hbp_isp.cu:981, 1012).  Both are integer algorithms whose
// results are fully determined by their inputs.
//
// Scan: tiles of 2048 ints (256 lanes x 8 consecutive items): the tiles' sums, one workgroup's
// scan of the sums (in chunks of 2048 with a running carry), then each tile's local scan plus
// its offset.
//
// Radix sort: 8-bit digits from bit 0 up to end_bit, one pass per digit on double buffers.  A
// pass: (1) each tile of 4096 keys (256 lanes x 16) counts its digits into hist[digit][tile]
// (digit-major); (2) the exclusive scan of hist gives every (digit, tile) its first output
// position; (3) each tile ranks its keys stably -- wavefront w holds keys [1024 w, 1024 w + 1024)
// of the tile, lane l key 64 r + l in round r, so (wave, round, lane) is the keys' order; a
// round finds the lanes with equal digits by eight ballots, and the lowest such lane advances
// the wavefront's counter of that digit by their number -- then adds the counts of the lower
// wavefronts and the tile's offset, and scatters.
#include "hbp_internal.h"

namespace hbp {

// ---------------------------------------------------------------------------------- scan
constexpr int SC_THREADS = 256, SC_ITEMS = 8, SC_TILE = SC_THREADS * SC_ITEMS;

// inclusive scan of one int per lane over the 256-lane workgroup; returns the exclusive value
// and the workgroup total
__device__ __forceinline__ int block_exclusive_scan(int v, int *lds_waves, int &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds_waves[wave] = x;
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < SC_THREADS / 64; ++w) {
        const int s = lds_waves[w];
        before += w < wave ? s : 0;
        total += s;
    }
    __syncthreads();   // lds_waves is reused by the caller's next call
    return before + x - v;
}

__global__ __launch_bounds__(SC_THREADS) void scan_sums_kernel(const int *__restrict__ in, long n,
                                                               int *__restrict__ sums) {
    __shared__ int w[SC_THREADS / 64];
    const long base = (long)blockIdx.x * SC_TILE + (long)threadIdx.x * SC_ITEMS;
    int s = 0;
#pragma unroll
    for (int i = 0; i < SC_ITEMS; ++i) s += base + i < n ? in[base + i] : 0;
    int total;
    (void)block_exclusive_scan(s, w, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// one workgroup: exclusive scan of the tiles' sums in place, chunk after chunk
__global__ __launch_bounds__(SC_THREADS) void scan_sums_scan_kernel(int *__restrict__ sums, int ntiles) {
    __shared__ int w[SC_THREADS / 64];
    int carry = 0;
    for (int c0 = 0; c0 < ntiles; c0 += SC_TILE) {
        const int base = c0 + threadIdx.x * SC_ITEMS;
        int v[SC_ITEMS], s = 0;
#pragma unroll
        for (int i = 0; i < SC_ITEMS; ++i) {
            v[i] = base + i < ntiles ? sums[base + i] : 0;
            s += v[i];
        }
        int total;
        int ex = block_exclusive_scan(s, w, total) + carry;
#pragma unroll
        for (int i = 0; i < SC_ITEMS; ++i) {
            if (base + i < ntiles) sums[base + i] = ex;
            ex += v[i];
        }
        carry += total;
    }
}

__global__ __launch_bounds__(SC_THREADS) void scan_apply_kernel(const int *__restrict__ in, long n,
                                                                const int *__restrict__ sums, int *__restrict__ out) {
    __shared__ int w[SC_THREADS / 64];
    const long base = (long)blockIdx.x * SC_TILE + (long)threadIdx.x * SC_ITEMS;
    int v[SC_ITEMS], s = 0;
#pragma unroll
    for (int i = 0; i < SC_ITEMS; ++i) {
        v[i] = base + i < n ? in[base + i] : 0;
        s += v[i];
    }
    int total;
    int ex = block_exclusive_scan(s, w, total) + (sums ? sums[blockIdx.x] : 0);
#pragma unroll
    for (int i = 0; i < SC_ITEMS; ++i) {
        if (base + i < n) out[base + i] = ex;
        ex += v[i];
    }
}

static long scan_tiles(long n) { return div_up(n, SC_TILE); }

// out[i] = in[0] + ... + in[i-1] (out may be in); sums: scan_tiles(n) ints of scratch
static void scan_exclusive_buf(Ctx &c, const int *d_in, int *d_out, long n, int *sums) {
    if (n <= 0) return;
    const long ntiles = scan_tiles(n);
    HBP_REQUIRE(ntiles < (1L << 30), "scan too long");
    if (ntiles == 1) {   // one tile: its local scan is the whole scan
        hipLaunchKernelGGL(scan_apply_kernel, dim3(1), dim3(SC_THREADS), 0, c.stream, d_in, n, (const int *)nullptr,
                           d_out);
        HBP_CHECK(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(scan_sums_kernel, dim3((unsigned)ntiles), dim3(SC_THREADS), 0, c.stream, d_in, n, sums);
    hipLaunchKernelGGL(scan_sums_scan_kernel, dim3(1), dim3(SC_THREADS), 0, c.stream, sums, (int)ntiles);
    hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)ntiles), dim3(SC_THREADS), 0, c.stream, d_in, n,
                       (const int *)sums, d_out);
    HBP_CHECK(hipGetLastError());
}

void scan_exclusive_int(Ctx &c, const int *d_in, int *d_out, long n, int *d_total) {
    (void)d_total;
    if (n <= 0) return;
    scan_exclusive_buf(c, d_in, d_out, n, (int *)c.scratch_bytes(sizeof(int) * (size_t)scan_tiles(n)));
}

// ---------------------------------------------------------------------------------- sort
constexpr int RS_THREADS = 256, RS_WAVES = RS_THREADS / 64, RS_ITEMS = 16, RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_BINS = 256;

template <class K>
__device__ __forceinline__ int digit_of(K k, int shift, unsigned mask) {
    return (int)((unsigned)(k >> shift) & mask);
}

template <class K>
__global__ __launch_bounds__(RS_THREADS) void rs_hist_kernel(const K *__restrict__ keys, long n, int shift,
                                                             unsigned mask, int ntiles, int *__restrict__ hist) {
    __shared__ int cnt[RS_WAVES][RS_BINS];
    for (int i = threadIdx.x; i < RS_WAVES * RS_BINS; i += RS_THREADS) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long base = (long)blockIdx.x * RS_TILE + (long)wave * 64 * RS_ITEMS + lane;
#pragma unroll 4
    for (int r = 0; r < RS_ITEMS; ++r) {
        const long i = base + (long)r * 64;
        if (i < n) atomicAdd(&cnt[wave][digit_of(keys[i], shift, mask)], 1);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < RS_BINS; d += RS_THREADS) {
        int s = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) s += cnt[w][d];
        hist[(long)d * ntiles + blockIdx.x] = s;
    }
}

// up to SC_TILE tiles: one workgroup per digit d scans that digit's row of tile counts,
// offs[d][t] = hist[d][0] + ... + hist[d][t-1], and leaves the row's total in totals[d]; the
// scatter adds the digits' bases (the exclusive scan of totals), which makes the offsets of the
// digit-major scan in one launch instead of three
__global__ __launch_bounds__(SC_THREADS) void rs_rowscan_kernel(const int *__restrict__ hist, int ntiles,
                                                              int *__restrict__ offs, int *__restrict__ totals) {
    __shared__ int w[SC_THREADS / 64];
    const long row = (long)blockIdx.x * ntiles;
    const int base = threadIdx.x * SC_ITEMS;
    int v[SC_ITEMS], s = 0;
#pragma unroll
    for (int i = 0; i < SC_ITEMS; ++i) {
        v[i] = base + i < ntiles ? hist[row + base + i] : 0;
        s += v[i];
    }
    int total;
    int ex = block_exclusive_scan(s, w, total);
#pragma unroll
    for (int i = 0; i < SC_ITEMS; ++i) {
        if (base + i < ntiles) offs[row + base + i] = ex;
        ex += v[i];
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

// totals: null when offs holds the digit-major scan itself, else the digits' totals
// (rs_rowscan_kernel; offs then holds each digit's row scan)
template <class K, class V, bool HAS_V>
__global__ __launch_bounds__(RS_THREADS) void rs_scatter_kernel(const K *__restrict__ kin, K *__restrict__ kout,
                                                                const V *__restrict__ vin, V *__restrict__ vout,
                                                                long n, int shift, unsigned mask, int ntiles,
                                                                const int *__restrict__ offs,
                                                                const int *__restrict__ totals) {
    static_assert(RS_BINS == RS_THREADS, "one digit per thread");
    __shared__ int wc[RS_WAVES][RS_BINS];   // per wavefront: running counts, then first positions
    __shared__ int dbase[RS_BINS];          // with totals: the keys of lower digits
    __shared__ int ws[RS_THREADS / 64];
    for (int i = threadIdx.x; i < RS_WAVES * RS_BINS; i += RS_THREADS) (&wc[0][0])[i] = 0;
    if (totals) {
        int all;
        dbase[threadIdx.x] = block_exclusive_scan(totals[threadIdx.x], ws, all);
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    const long base = (long)blockIdx.x * RS_TILE + (long)wave * 64 * RS_ITEMS + lane;
    K k[RS_ITEMS];
    V v[RS_ITEMS];
    int rank[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const long i = base + (long)r * 64;
        const bool ok = i < n;
        k[r] = ok ? kin[i] : (K)0;
        if (HAS_V) v[r] = ok ? vin[i] : (V)0;
    }
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const bool ok = base + (long)r * 64 < n;
        const int d = digit_of(k[r], shift, mask);
        unsigned long long peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long bal = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? bal : ~bal;
        }
        int old = 0;
        const int leader = __ffsll((long long)peers) - 1;   // lanes with an invalid key have no peers
        if (ok && lane == leader) {
            old = wc[wave][d];
            wc[wave][d] = old + __popcll(peers);
        }
        old = __shfl(old, leader < 0 ? lane : leader, 64);
        rank[r] = old + __popcll(peers & below);
    }
    __syncthreads();
    // first position of each (wavefront, digit): the tile's offset plus the lower wavefronts' counts
    for (int d = threadIdx.x; d < RS_BINS; d += RS_THREADS) {
        int pos = offs[(long)d * ntiles + blockIdx.x] + (totals ? dbase[d] : 0);
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) {
            const int cnt = wc[w][d];
            wc[w][d] = pos;
            pos += cnt;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        if (base + (long)r * 64 < n) {
            const int p = wc[wave][digit_of(k[r], shift, mask)] + rank[r];
            kout[p] = k[r];
            if (HAS_V) vout[p] = v[r];
        }
    }
}

// every pass of a sort of n <= RS_TILE keys in one workgroup: the keys (and values) stay in
// registers and LDS; the same stable ranking as rs_scatter_kernel, the digits' offsets from a
// workgroup scan of their counts; the result goes back to keys / vals
template <class K, class V, bool HAS_V>
__global__ __launch_bounds__(RS_THREADS) void rs_block_sort_kernel(K *__restrict__ keys, V *__restrict__ vals, int n,
                                                                   int end_bit) {
    __shared__ int wc[RS_WAVES][RS_BINS];
    __shared__ int wsum[RS_THREADS / 64];
    __shared__ K lk[RS_TILE];
    __shared__ V lv[HAS_V ? RS_TILE : 1];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int base = wave * 64 * RS_ITEMS + lane;
    K k[RS_ITEMS];
    V v[RS_ITEMS];
    int rank[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int i = base + r * 64;
        k[r] = i < n ? keys[i] : (K)0;
        if (HAS_V) v[r] = i < n ? vals[i] : (V)0;
    }
    for (int shift = 0; shift < end_bit; shift += 8) {
        const int bits = end_bit - shift < 8 ? end_bit - shift : 8;
        const unsigned mask = (1u << bits) - 1u;
        for (int i = threadIdx.x; i < RS_WAVES * RS_BINS; i += RS_THREADS) (&wc[0][0])[i] = 0;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RS_ITEMS; ++r) {
            const bool ok = base + r * 64 < n;
            const int d = digit_of(k[r], shift, mask);
            unsigned long long peers = __ballot(ok);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const unsigned long long bal = __ballot((d >> b) & 1);
                peers &= ((d >> b) & 1) ? bal : ~bal;
            }
            int old = 0;
            const int leader = __ffsll((long long)peers) - 1;
            if (ok && lane == leader) {
                old = wc[wave][d];
                wc[wave][d] = old + __popcll(peers);
            }
            old = __shfl(old, leader < 0 ? lane : leader, 64);
            rank[r] = old + __popcll(peers & below);
        }
        __syncthreads();
        {   // thread d: the digit's first position, then each wavefront's
            const int d = threadIdx.x;
            int cnt = 0;
#pragma unroll
            for (int w = 0; w < RS_WAVES; ++w) cnt += wc[w][d];
            int total;
            int pos = block_exclusive_scan(cnt, wsum, total);
#pragma unroll
            for (int w = 0; w < RS_WAVES; ++w) {
                const int c = wc[w][d];
                wc[w][d] = pos;
                pos += c;
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RS_ITEMS; ++r) {
            if (base + r * 64 < n) {
                const int p = wc[wave][digit_of(k[r], shift, mask)] + rank[r];
                lk[p] = k[r];
                if (HAS_V) lv[p] = v[r];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RS_ITEMS; ++r) {
            const int i = base + r * 64;
            k[r] = i < n ? lk[i] : (K)0;
            if (HAS_V) v[r] = i < n ? lv[i] : (V)0;
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int i = base + r * 64;
        if (i < n) {
            keys[i] = k[r];
            if (HAS_V) vals[i] = v[r];
        }
    }
}

// Stable LSD radix sort of n keys on bits [0, end_bit) (bits above end_bit must be zero), the
// values permuted with them; double buffers: returns true when the result is in the *_tmp
// buffers (the caller swaps its buffers), which saves a copy back.
template <class K, class V, bool HAS_V>
static bool radix_sort(Ctx &c, K *keys, K *keys_tmp, V *vals, V *vals_tmp, long n, int end_bit) {
    if (n <= 1 || end_bit <= 0) return false;
    HBP_REQUIRE(n < (1L << 31), "sort too long");
    if (n <= RS_TILE) {   // one workgroup, every pass, in place
        hipLaunchKernelGGL((rs_block_sort_kernel<K, V, HAS_V>), dim3(1), dim3(RS_THREADS), 0, c.stream, keys, vals,
                           (int)n, end_bit);
        HBP_CHECK(hipGetLastError());
        return false;
    }
    const long ntiles = div_up(n, RS_TILE);
    // scratch: hist [BINS][tiles], its exclusive scan (or each digit's row scan), the scan's tile
    // sums (or the digits' totals)
    const long nh = (long)RS_BINS * ntiles;
    const bool rows = ntiles <= SC_TILE;
    int *hist = (int *)c.scratch_bytes(sizeof(int) * (size_t)(2 * nh + std::max<long>(scan_tiles(nh), RS_BINS)));
    int *offs = hist + nh, *sums = offs + nh;
    K *ka = keys, *kb = keys_tmp;
    V *va = vals, *vb = vals_tmp;
    bool flipped = false;
    for (int shift = 0; shift < end_bit; shift += 8) {
        const int bits = end_bit - shift < 8 ? end_bit - shift : 8;
        const unsigned mask = (1u << bits) - 1u;
        hipLaunchKernelGGL(rs_hist_kernel<K>, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, c.stream, ka, n, shift,
                           mask, (int)ntiles, hist);
        if (rows)
            hipLaunchKernelGGL(rs_rowscan_kernel, dim3(RS_BINS), dim3(SC_THREADS), 0, c.stream, (const int *)hist,
                               (int)ntiles, offs, sums);
        else
            scan_exclusive_buf(c, hist, offs, nh, sums);
        hipLaunchKernelGGL((rs_scatter_kernel<K, V, HAS_V>), dim3((unsigned)ntiles), dim3(RS_THREADS), 0, c.stream, ka,
                           kb, va, vb, n, shift, mask, (int)ntiles, (const int *)offs,
                           rows ? (const int *)sums : (const int *)nullptr);
        HBP_CHECK(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
        flipped = !flipped;
    }
    return flipped;
}

bool sort_pairs_u64(Ctx &c, unsigned long long *keys, unsigned long long *keys_tmp, unsigned int *vals,
                    unsigned int *vals_tmp, long n, int end_bit) {
    return radix_sort<unsigned long long, unsigned int, true>(c, keys, keys_tmp, vals, vals_tmp, n, end_bit);
}

bool sort_pairs_u32(Ctx &c, unsigned int *keys, unsigned int *keys_tmp, int *vals, int *vals_tmp, long n,
                    int end_bit) {
    return radix_sort<unsigned int, int, true>(c, keys, keys_tmp, vals, vals_tmp, n, end_bit);
}

bool sort_keys_u64(Ctx &c, unsigned long long *keys, unsigned long long *keys_tmp, long n, int end_bit) {
    return radix_sort<unsigned long long, unsigned int, false>(c, keys, keys_tmp, (unsigned int *)nullptr,
                                                               (unsigned int *)nullptr, n, end_bit);
}

}  // namespace hbp
