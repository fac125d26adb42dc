// radix_sort.hip - stable LSD radix sort of (cell-hash key, particle index) pairs
// for gfx950.  Replaces the reference's clogs::Radixsort call
// (PCISPH/src/sim/Fluid.cpp:84-87,153,193), whose contract is a stable key/value
// sort (libs/clogs/include/clogs/radixsort.h:182-184).  Written from that contract
// and the published reduce-then-scan radix sort (Merrill & Grimshaw 2011).
//
// Per digit pass:
//   upsweep   : each tile of 2048 keys builds its digit histogram in LDS and writes
//               it digit-major, hist[d * tiles + t]
//   scan      : hist becomes exclusive offsets within chunks of 2048 words, the
//               chunk totals beside them (one launch); with at most kFusedChunks
//               chunks each downsweep workgroup scans the totals itself and adds them
//               back (else two more launches finish the scan first)
//   downsweep : each tile re-reads its keys, ranks them stably with wavefront
//               ballots (the lanes sharing a digit are found with one ballot per
//               digit bit; a lane's rank is the popcount of those below it), adds
//               the per-wavefront prefix kept in LDS, and scatters key and value.
// Digit width is a runtime choice of 4..8 bits.
#include <hip/hip_runtime.h>

#include "sph_launch.h"

namespace pcisph {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxRadix = 256;

__global__ __launch_bounds__(kThreads) void k_upsweep(const uint32_t* __restrict__ keys, uint32_t n,
                                                       uint32_t shift, uint32_t bits,
                                                       uint32_t* __restrict__ hist, uint32_t tiles) {
    __shared__ uint32_t s_hist[kMaxRadix];
    const uint32_t radix = 1u << bits, mask = radix - 1;
    for (uint32_t d = threadIdx.x; d < radix; d += kThreads) s_hist[d] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kTile;
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const uint32_t e = base + r * kThreads + threadIdx.x;
        if (e < n) atomicAdd(&s_hist[(keys[e] >> shift) & mask], 1u);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < radix; d += kThreads) hist[d * tiles + blockIdx.x] = s_hist[d];
}

// Exclusive scan of `count` words: chunks of kScanChunk words scanned by one
// workgroup each (k_scan_chunks, chunk totals to `partials`), the totals scanned --
// by one workgroup (k_scan_totals) when there are at most 1024 of them, else by the
// same scan one level up -- and added back (k_scan_add).
constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr int kScanChunk = kScanThreads * kScanPer;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_tmp[wave] = x;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        if (w < wave) before += s_tmp[w];
        total += s_tmp[w];
    }
    __syncthreads();
    return before + x - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_chunks(uint32_t* __restrict__ data, uint32_t count,
                                                              uint32_t* __restrict__ partials) {
    __shared__ uint32_t s_tmp[kScanThreads / 64];
    const uint32_t base = blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
    uint32_t v[kScanPer], sum = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        v[k] = base + k < count ? data[base + k] : 0u;
        sum += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, s_tmp, total);
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        if (base + k < count) data[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_totals(uint32_t* __restrict__ partials, uint32_t chunks) {
    __shared__ uint32_t s_tmp[16];
    const uint32_t v = threadIdx.x < chunks ? partials[threadIdx.x] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(v, s_tmp, total);
    if (threadIdx.x < chunks) partials[threadIdx.x] = ex;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(uint32_t* __restrict__ data, uint32_t count,
                                                           const uint32_t* __restrict__ partials) {
    const uint32_t add = partials[blockIdx.x];
    const uint32_t base = blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k)
        if (base + k < count) data[base + k] += add;
}

// Chunk totals a downsweep workgroup scans itself (8 per thread).
constexpr uint32_t kFusedChunks = 2048;

// `offsets`: exclusive global offsets, or (with `partials`) offsets within the scan's
// chunks, `partials` holding the `chunks` chunk totals.  The tile is ranked whole
// before anything is written: every item's rank among the lanes of its wavefront that
// hold the same digit (one ballot per digit bit), digit counts per (round, wavefront),
// their scan in element order per digit and over the digits, then the tile in its
// stable sorted order in LDS, written out in runs of one digit (consecutive threads,
// consecutive global slots).
__global__ __launch_bounds__(kThreads) void k_downsweep(const uint32_t* __restrict__ kin,
                                                         const uint32_t* __restrict__ vin,
                                                         uint32_t* __restrict__ kout,
                                                         uint32_t* __restrict__ vout, uint32_t n,
                                                         uint32_t shift, uint32_t bits,
                                                         const uint32_t* __restrict__ offsets,
                                                         uint32_t tiles, const uint32_t* __restrict__ partials,
                                                         uint32_t chunks) {
    __shared__ uint32_t s_run[kMaxRadix];                   // global slot of each digit's first item
    __shared__ uint32_t s_start[kMaxRadix];                 // tile-local slot of each digit's first item
    __shared__ uint32_t s_cnt[kItems * kWaves][kMaxRadix];  // digit counts per (round, wavefront), then bases
    __shared__ uint32_t s_key[kTile], s_val[kTile];         // the tile in sorted order
    __shared__ uint32_t s_pre[kFusedChunks];                // exclusive prefix of the chunk totals
    __shared__ uint32_t s_tmp[kWaves];
    const uint32_t radix = 1u << bits, mask = radix - 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (partials) {
        constexpr uint32_t per = kFusedChunks / kThreads;
        const uint32_t c0 = threadIdx.x * per;
        uint32_t v[per], sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < per; ++k) {
            v[k] = c0 + k < chunks ? partials[c0 + k] : 0u;
            sum += v[k];
        }
        uint32_t total;
        uint32_t run = block_exclusive_scan(sum, s_tmp, total);
#pragma unroll
        for (uint32_t k = 0; k < per; ++k) {
            s_pre[c0 + k] = run;
            run += v[k];
        }
        __syncthreads();
    }
    for (uint32_t d = threadIdx.x; d < radix; d += kThreads) {
        const uint32_t w = d * tiles + blockIdx.x;
        s_run[d] = offsets[w] + (partials ? s_pre[w / kScanChunk] : 0u);
#pragma unroll 8
        for (int k = 0; k < kItems * kWaves; ++k) s_cnt[k][d] = 0;
    }
    __syncthreads();
    // 1. ranks within the wavefront, counts per (round, wavefront, digit)
    const unsigned long long lt = (1ull << lane) - 1ull;
    const uint32_t base = blockIdx.x * kTile;
    uint32_t key[kItems], val[kItems], rank[kItems];
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const uint32_t e = base + r * kThreads + threadIdx.x;
        const bool valid = e < n;
        key[r] = valid ? kin[e] : 0u;
        val[r] = valid ? vin[e] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const bool valid = base + r * kThreads + threadIdx.x < n;
        const uint32_t digit = (key[r] >> shift) & mask;
        unsigned long long same = __ballot(valid);
        for (uint32_t b = 0; b < bits; ++b) {
            const bool bit = (digit >> b) & 1u;
            const unsigned long long set = __ballot(bit);
            same &= bit ? set : ~set;
        }
        rank[r] = (uint32_t)__popcll(same & lt);
        if (valid && rank[r] == 0) s_cnt[r * kWaves + wave][digit] = (uint32_t)__popcll(same);
    }
    __syncthreads();
    // 2. per digit, the counts scanned in element order (round, then wavefront); then
    // the digits' totals scanned into tile-local starts
    uint32_t total = 0;
    if (threadIdx.x < radix) {
        const uint32_t d = threadIdx.x;
#pragma unroll 4
        for (int k = 0; k < kItems * kWaves; ++k) {
            const uint32_t c = s_cnt[k][d];
            s_cnt[k][d] = total;
            total += c;
        }
    }
    uint32_t all;
    const uint32_t start = block_exclusive_scan(total, s_tmp, all);
    if (threadIdx.x < radix) s_start[threadIdx.x] = start;
    __syncthreads();
    // 3. the tile in sorted order in LDS
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        if (base + r * kThreads + threadIdx.x < n) {
            const uint32_t digit = (key[r] >> shift) & mask;
            const uint32_t at = s_start[digit] + s_cnt[r * kWaves + wave][digit] + rank[r];
            s_key[at] = key[r];
            s_val[at] = val[r];
        }
    }
    __syncthreads();
    // 4. runs of one digit to consecutive global slots
    const uint32_t count = n - base < (uint32_t)kTile ? n - base : (uint32_t)kTile;
    for (uint32_t i = threadIdx.x; i < count; i += kThreads) {
        const uint32_t k = s_key[i], digit = (k >> shift) & mask;
        const uint32_t dst = s_run[digit] + (i - s_start[digit]);
        kout[dst] = k;
        vout[dst] = s_val[i];
    }
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* p, size_t n, unsigned long long v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

void fill_u32(const Launch& L, uint32_t* p, size_t n, uint32_t value) {
    if (!n) return;
    const uint32_t blocks = (uint32_t)((n + 255) / 256);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, L.stream, p, n, value);
}

void fill_u64(const Launch& L, unsigned long long* p, size_t n, unsigned long long value) {
    if (!n) return;
    const uint32_t blocks = (uint32_t)((n + 255) / 256);
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, L.stream, p, n, value);
}

// Words of chunk totals the scan of `count` words needs at every level.
size_t scan_partial_words(size_t count) {
    size_t words = 0;
    for (size_t c = (count + kScanChunk - 1) / kScanChunk; c > 1; c = (c + kScanChunk - 1) / kScanChunk)
        words += c;
    return words + 1024;
}

// Exclusive scan of `count` words.  With `leave_chunks`, the scan stops after the
// chunk scans when there are at most kFusedChunks chunks and returns their count
// (`data` then holds offsets within chunks, `partials` the chunk totals); otherwise
// it completes and returns 0.
uint32_t exclusive_scan(const Launch& L, uint32_t* data, uint32_t count, uint32_t* partials,
                        bool leave_chunks = false) {
    const uint32_t chunks = (count + kScanChunk - 1) / kScanChunk;
    hipLaunchKernelGGL(k_scan_chunks, dim3(chunks), dim3(kScanThreads), 0, L.stream, data, count, partials);
    if (leave_chunks && chunks <= kFusedChunks) return chunks;
    if (chunks <= 1) return 0;
    if (chunks <= 1024) hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(1024), 0, L.stream, partials, chunks);
    else exclusive_scan(L, partials, chunks, partials + chunks);
    hipLaunchKernelGGL(k_scan_add, dim3(chunks), dim3(kScanThreads), 0, L.stream, data, count, partials);
    return 0;
}

size_t radix_sort_scratch_words(uint32_t n, uint32_t digit_bits) {
    const size_t tiles = (n + kTile - 1) / kTile;
    const size_t words = (size_t(1) << digit_bits) * (tiles ? tiles : 1);
    const size_t own = words + scan_partial_words(words);  // histograms, then the scan's chunk totals
    const size_t one = radix_onesweep_scratch_words(n, digit_bits);
    return own > one ? own : one;
}

int radix_sort_result(uint32_t n, uint32_t key_bits, uint32_t digit_bits) {
    if (n <= 1 || key_bits == 0) return 0;
    return (int)(((key_bits + digit_bits - 1) / digit_bits) & 1u);
}

int radix_sort(const Launch& L, uint32_t* keys[2], uint32_t* vals[2], uint32_t n, uint32_t key_bits,
               uint32_t digit_bits, uint32_t* scratch) {
    if (L.onesweep) return radix_sort_onesweep(L, keys, vals, n, key_bits, digit_bits, scratch);
    if (n <= 1 || key_bits == 0) return 0;
    const uint32_t tiles = (n + kTile - 1) / kTile;
    int cur = 0;
    for (uint32_t shift = 0; shift < key_bits; shift += digit_bits) {
        const uint32_t bits = key_bits - shift < digit_bits ? key_bits - shift : digit_bits;
        const uint32_t words = (1u << bits) * tiles;
        hipLaunchKernelGGL(k_upsweep, dim3(tiles), dim3(kThreads), 0, L.stream, keys[cur], n, shift, bits,
                           scratch, tiles);
        uint32_t* partials = scratch + (size_t(1) << digit_bits) * tiles;
        const uint32_t chunks = exclusive_scan(L, scratch, words, partials, true);
        hipLaunchKernelGGL(k_downsweep, dim3(tiles), dim3(kThreads), 0, L.stream, keys[cur], vals[cur],
                           keys[cur ^ 1], vals[cur ^ 1], n, shift, bits, scratch, tiles,
                           chunks ? partials : nullptr, chunks);
        cur ^= 1;
    }
    return cur;
}

}  // namespace pcisph
