// radix_onesweep.hip - the one-sweep variant of the stable LSD radix sort of
// (cell-hash key, particle index) pairs, selected by solver option work_order bit 7
// (radix_sort.hip holds the default, reduce-then-scan).  Same contract as the
// reference's clogs::Radixsort call (PCISPH/src/sim/Fluid.cpp:84-87,153,193; stable
// key/value sort, libs/clogs/include/clogs/radixsort.h:182-184), written from the
// published one-sweep radix sort (Adinets & Merrill 2022): one scatter pass per digit,
// its global offsets from a decoupled look-back over tile prefixes.
//
//   k_onesweep_hist : one read of the keys builds every pass's digit histogram (LDS,
//                     then device atomics) and clears the tiles' look-back words
//   k_onesweep_scan : the histograms become each pass's exclusive digit bases
//   k_onesweep_pass : per digit pass, each workgroup takes the next tile in launch
//                     order (a device counter, so the tiles it waits on are held by
//                     workgroups already running), ranks its keys stably with
//                     wavefront ballots (the lanes sharing a digit are found with one
//                     ballot per digit bit; a lane's rank is the popcount of those
//                     below it) and per-(item, wavefront) digit counts in LDS,
//                     publishes its digit totals, adds up its predecessors' totals
//                     back to the first inclusive prefix found, publishes its own
//                     inclusive prefix and scatters key and value.
// Digit width is a runtime choice of 4..8 bits.  The look-back wait is bounded in
// wall-clock time: past the bound it sets an abort word, raises the caller's report
// flag (host-visible) and stops waiting; the destination slots it then computes are
// below the true ones, so every store stays in bounds, and the step reports the
// failure.  Measured slower than reduce-then-scan on MI355X (the look-back is a
// latency chain through the resident tiles), hence not the default.
#include <hip/hip_runtime.h>

#include "sph_launch.h"

namespace pcisph {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxRadix = 256;
constexpr int kMaxPasses = 8;  // 32-bit keys at 4-bit digits
// look-back words: flag in the high half, count or prefix in the low half
constexpr unsigned long long kAggregate = 1ull << 32;  // the tile's own digit total
constexpr unsigned long long kPrefix = 2ull << 32;     // inclusive of every earlier tile

// Scratch layout (words): [0, kMaxPasses * kMaxRadix) digit histograms, then bases;
// kMaxPasses tile counters; abort word; pad to 8 bytes; look-back words
// [passes][tiles][radix] (64-bit).
constexpr size_t kHistWords = (size_t)kMaxPasses * kMaxRadix;
constexpr size_t kCounterWord = kHistWords;
constexpr size_t kAbortWord = kCounterWord + kMaxPasses;
constexpr size_t kStatusWord = (kAbortWord + 2) & ~size_t(1);

__global__ __launch_bounds__(kThreads) void k_onesweep_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                             uint32_t key_bits, uint32_t bits, uint32_t passes,
                                                             uint32_t* __restrict__ hist,
                                                             unsigned long long* __restrict__ status,
                                                             uint32_t tiles) {
    __shared__ uint32_t s_hist[kMaxPasses][kMaxRadix];
    const uint32_t radix = 1u << bits, mask = radix - 1;
    for (uint32_t q = threadIdx.x; q < passes * radix; q += kThreads) s_hist[q / radix][q % radix] = 0;
    // this tile's look-back words of every pass start cleared
    for (uint32_t q = threadIdx.x; q < passes * radix; q += kThreads)
        status[((size_t)(q / radix) * tiles + blockIdx.x) * radix + q % radix] = 0ull;
    __syncthreads();
    const uint32_t base = blockIdx.x * kTile;
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const uint32_t e = base + r * kThreads + threadIdx.x;
        if (e >= n) continue;
        const uint32_t key = keys[e];
        for (uint32_t p = 0; p < passes; ++p) atomicAdd(&s_hist[p][(key >> (p * bits)) & mask], 1u);
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < passes * radix; q += kThreads) {
        const uint32_t c = s_hist[q / radix][q % radix];
        if (c) atomicAdd(&hist[(q / radix) * kMaxRadix + q % radix], c);
    }
    (void)key_bits;
}

// One workgroup: each pass's histogram into exclusive digit bases, in place.
__global__ __launch_bounds__(kMaxRadix) void k_onesweep_scan(uint32_t* __restrict__ hist, uint32_t bits,
                                                             uint32_t passes) {
    __shared__ uint32_t s_tmp[kMaxRadix / 64];
    const uint32_t radix = 1u << bits;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t v = threadIdx.x < radix ? hist[p * kMaxRadix + threadIdx.x] : 0u;
        uint32_t x = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) s_tmp[wave] = x;
        __syncthreads();
        uint32_t before = 0;
        for (int w = 0; w < wave; ++w) before += s_tmp[w];
        if (threadIdx.x < radix) hist[p * kMaxRadix + threadIdx.x] = before + x - v;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kThreads) void k_onesweep_pass(const uint32_t* __restrict__ kin,
                                                             const uint32_t* __restrict__ vin,
                                                             uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                             uint32_t n, uint32_t shift, uint32_t bits,
                                                             const uint32_t* __restrict__ bases,
                                                             uint32_t* __restrict__ counter,
                                                             unsigned long long* __restrict__ status,
                                                             uint32_t* __restrict__ abort_word,
                                                             unsigned int* __restrict__ report,
                                                             unsigned long long timeout) {
    __shared__ uint32_t s_cnt[kItems][kWaves][kMaxRadix];  // per (item, wavefront) digit count, then offset
    __shared__ uint32_t s_glob[kMaxRadix];                 // the tile's first global slot per digit
    __shared__ uint32_t s_tile;
    const uint32_t radix = 1u << bits, mask = radix - 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
    for (uint32_t q = threadIdx.x; q < kItems * kWaves * radix; q += kThreads)
        s_cnt[q / (kWaves * radix)][(q / radix) % kWaves][q % radix] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * kTile;
    const unsigned long long lt = (1ull << lane) - 1ull;
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
        const uint32_t e = base + r * kThreads + threadIdx.x;
        const bool valid = e < n;
        const uint32_t digit = (key[r] >> shift) & mask;
        unsigned long long same = __ballot(valid);
        for (uint32_t b = 0; b < bits; ++b) {
            const bool bit = (digit >> b) & 1u;
            const unsigned long long set = __ballot(bit);
            same &= bit ? set : ~set;
        }
        rank[r] = (uint32_t)__popcll(same & lt);
        if (valid && rank[r] == 0) s_cnt[r][wave][digit] = (uint32_t)__popcll(same);
    }
    __syncthreads();
    // per digit: offsets of (item, wavefront) groups in index order, the tile total,
    // then the decoupled look-back over the earlier tiles
    if (threadIdx.x < radix) {
        const uint32_t d = threadIdx.x;
        uint32_t run = 0;
#pragma unroll
        for (int r = 0; r < kItems; ++r)
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t c = s_cnt[r][w][d];
                s_cnt[r][w][d] = run;
                run += c;
            }
        unsigned long long* const mine = status + (size_t)tile * radix + d;
        __hip_atomic_store(mine, (tile ? kAggregate : kPrefix) | run, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t prefix = 0;
        if (tile) {
            const unsigned long long t0 = (unsigned long long)wall_clock64();
            bool stop = false;
            for (uint32_t j = tile; j > 0 && !stop;) {
                const unsigned long long w =
                    __hip_atomic_load(status + (size_t)(j - 1) * radix + d, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                if (w >> 32) {
                    prefix += (uint32_t)w;
                    stop = (w & kPrefix) != 0;
                    --j;
                } else if (__hip_atomic_load(abort_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                           (unsigned long long)wall_clock64() - t0 > timeout) {
                    __hip_atomic_store(abort_word, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (report) __hip_atomic_store(report, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    stop = true;
                } else {
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __hip_atomic_store(mine, kPrefix | (prefix + run), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_glob[d] = bases[d] + prefix;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const uint32_t e = base + r * kThreads + threadIdx.x;
        if (e >= n) continue;
        const uint32_t digit = (key[r] >> shift) & mask;
        const uint32_t dst = s_glob[digit] + s_cnt[r][wave][digit] + rank[r];
        kout[dst] = key[r];
        vout[dst] = val[r];
    }
}

}  // namespace

size_t radix_onesweep_scratch_words(uint32_t n, uint32_t digit_bits) {
    const size_t tiles = (n + kTile - 1) / kTile;
    const size_t bits = digit_bits < 4 ? 4 : (digit_bits > 8 ? 8 : digit_bits);
    const size_t passes = (32 + bits - 1) / bits;
    return kStatusWord + 2 * passes * (tiles ? tiles : 1) * (size_t(1) << bits);
}

int radix_sort_onesweep(const Launch& L, uint32_t* keys[2], uint32_t* vals[2], uint32_t n, uint32_t key_bits,
                        uint32_t digit_bits, uint32_t* scratch) {
    if (n <= 1 || key_bits == 0) return 0;
    if (digit_bits < 4 || digit_bits > 8 || key_bits > 32) return -1;
    const uint32_t tiles = (n + kTile - 1) / kTile;
    const uint32_t passes = (key_bits + digit_bits - 1) / digit_bits;
    uint32_t* hist = scratch;
    unsigned long long* status = reinterpret_cast<unsigned long long*>(scratch + kStatusWord);
    // histograms, tile counters and the abort word start at zero
    fill_u32(L, scratch, kStatusWord, 0u);
    hipLaunchKernelGGL(k_onesweep_hist, dim3(tiles), dim3(kThreads), 0, L.stream, keys[0], n, key_bits, digit_bits,
                       passes, hist, status, tiles);
    hipLaunchKernelGGL(k_onesweep_scan, dim3(1), dim3(kMaxRadix), 0, L.stream, hist, digit_bits, passes);
    // the look-back gives up after 2 s of waiting (100 MHz wall clock)
    const unsigned long long timeout = 200000000ull;
    const size_t per_pass = (size_t)tiles << digit_bits;
    int cur = 0;
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = p * digit_bits;
        const uint32_t bits = key_bits - shift < digit_bits ? key_bits - shift : digit_bits;
        hipLaunchKernelGGL(k_onesweep_pass, dim3(tiles), dim3(kThreads), 0, L.stream, keys[cur], vals[cur],
                           keys[cur ^ 1], vals[cur ^ 1], n, shift, bits, hist + (size_t)p * kMaxRadix,
                           scratch + kCounterWord + p, status + p * per_pass, scratch + kAbortWord, L.sort_abort,
                           timeout);
        cur ^= 1;
    }
    return cur;
}

}  // namespace pcisph
