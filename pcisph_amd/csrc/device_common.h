// device_common.h - grid hashing, parameter blocks and vector helpers shared by
// every gfx950 kernel of the step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcisph {

template <typename R>
struct Vec4Of;
template <>
struct Vec4Of<float> {
    using type = float4;
};
template <>
struct Vec4Of<double> {
    using type = double4;
};
template <typename R>
using V4 = typename Vec4Of<R>::type;

// Bucket count plus what `key % bucket_count` needs without a hardware divide:
// a mask when the count is a power of two, else Lemire's direct-remainder constant
// M = floor((2^64-1)/B)+1, exact for every 32-bit dividend ("Faster Remainder by
// Direct Computation", Lemire, Kaser & Kurz 2019).
struct HashGrid {
    uint32_t bucket_count;
    uint32_t pow2_mask;  // bucket_count-1 if a power of two, else 0
    uint64_t fastmod_m;
};

inline HashGrid make_hash_grid(uint32_t B) {
    HashGrid g;
    g.bucket_count = B;
    g.pow2_mask = (B & (B - 1)) == 0 ? B - 1 : 0;
    g.fastmod_m = UINT64_MAX / B + 1;
    return g;
}

__device__ __forceinline__ uint32_t bucket_mod(uint32_t v, const HashGrid& g) {
    if (g.pow2_mask) return v & g.pow2_mask;
    const uint64_t low = g.fastmod_m * (uint64_t)v;
    return (uint32_t)__umul64hi(low, (uint64_t)g.bucket_count);
}

// get_hash_key (PCISPH/data/kernels/grid_utils.cl:12-28), 32-bit wrap arithmetic.
__device__ __forceinline__ uint32_t hash_key(uint32_t x, uint32_t y, uint32_t z, const HashGrid& g) {
    const uint32_t h = (x * 19349663u) ^ (y * 73856093u) ^ (z * 83492791u);
    const uint32_t v = (h << 6) + (x & 63u) + 4u * (y & 63u) + 16u * (z & 63u);
    return bucket_mod(v & 0x7FFFFFFFu, g);
}

// convert_int_sat_rtn (grid_utils.cl:9-11): floor, saturate, NaN -> 0.
template <typename R>
__device__ __forceinline__ int32_t sat_floor(R v) {
    if (v != v) return 0;
    const R f = floor(v);
    if (f >= (R)2147483648.0) return INT32_MAX;
    if (f < (R)-2147483648.0) return INT32_MIN;
    return (int32_t)f;
}

__device__ __forceinline__ float div_rn(float a, float b) { return a / b; }  // correctly rounded
__device__ __forceinline__ double div_rn(double a, double b) { return a / b; }

// The cell of a position: IEEE division in every mode, so keys are bit-exact.
template <typename R>
__device__ __forceinline__ int3 cell_of(R x, R y, R z, R cell) {
    return make_int3(sat_floor<R>(div_rn(x, cell)), sat_floor<R>(div_rn(y, cell)),
                     sat_floor<R>(div_rn(z, cell)));
}

// The 27 buckets around cell c in the reference order: offset index
// i = (dx+1) + 3(dy+1) + 9(dz+1), x fastest (grid_utils.cl:65-79).  The nine
// (dz, dy) rows are a loop, not unrolled: `fn` is a whole gather, and 27 inlined
// copies of it per walk make kernels of tens of thousands of instructions that
// thrash the instruction cache.  The three x buckets of a row are unrolled, with
// the x hash products formed once.
template <typename F>
__device__ __forceinline__ void for_each_bucket_row(int3 c, const HashGrid& g, F&& fn) {
    uint32_t hx[3], lx[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const uint32_t x = (uint32_t)c.x + (uint32_t)(d - 1);
        hx[d] = x * 19349663u;
        lx[d] = x & 63u;
    }
#pragma unroll 1
    for (int row = 0; row < 9; ++row) {
        const uint32_t y = (uint32_t)c.y + (uint32_t)(row % 3 - 1);
        const uint32_t z = (uint32_t)c.z + (uint32_t)(row / 3 - 1);
        const uint32_t hyz = (y * 73856093u) ^ (z * 83492791u);
        const uint32_t lyz = 4u * (y & 63u) + 16u * (z & 63u);
        uint32_t k[3];
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) k[dx] = bucket_mod((((hx[dx] ^ hyz) << 6) + lx[dx] + lyz) & 0x7FFFFFFFu, g);
        fn(k[0], k[1], k[2]);
    }
}

// Where a scene's bucket ranges are read.  `cells` is the scene's slice of the
// bucket table (indexed by hash key).  `grid`, when dim[0] != 0, is a dense copy for
// the cells of a box [lo, lo + dim): entry rx + dim[0] (ry + dim[1] rz) holds the
// range of the bucket that cell lo + (rx, ry, rz) hashes to, so a neighbourhood's
// lookups read neighbouring words instead of hash-scattered ones.  Both give the
// same range for a cell; cells outside the box use the bucket table.
//
// A grid entry may carry kFarBucket in its start word: every particle of that bucket
// lies in a cell at least kFarCells cells from the entry's cell along some axis (hash
// collisions from elsewhere in the scene), hence at least kFarCells - 2 - 0.01 cell
// sizes from every position whose cell is the entry's cell or one of its 26
// neighbours -- the positions that look the entry up (the 0.01 covers the rounding of
// cell coordinates of magnitude below 2^15, the only ones flagged).  A walk whose
// evaluation points stay within d of the walking particle's position, and whose terms
// vanish beyond radius rho, adds exactly zero for such a bucket when d + rho <=
// kFarReach cell sizes: it sets `prune` and reads those entries as empty.  Candidate
// counts still include them.  Eight cells keep a walk prunable while its particle's
// predicted position is up to ~5 cells from where the step began (the start-up
// transient of the dam breaks moves 6-20% of the particles by more than one cell
// within a step), and hash collisions from closer than that are rare.
//
// kQuietCell in an entry's end word marks a cell all 27 of whose neighbourhood
// entries are empty or far: a pruning walk from it has nothing to visit.
constexpr uint32_t kFarBucket = 0x80000000u;
constexpr uint32_t kQuietCell = 0x80000000u;
constexpr int kFarCells = 8;
constexpr float kFarReach = (float)kFarCells - 2.02f;

struct CellTable {
    const uint2* cells;
    const uint2* grid;
    HashGrid g;
    uint32_t lo[3];   // box origin (cell coordinates as 32-bit words)
    uint32_t dim[3];  // box size in cells; dim[0] == 0: no box
    bool prune;       // grid entries flagged kFarBucket read as empty
};

// The bucket ranges of the 27 cells around c, one (dz, dy) row of three x-adjacent
// cells at a time, in the reference's order (grid_utils.cl:65-79): fn(r0, r1, r2).
// A row inside the box reads three consecutive grid entries; any other row hashes.
// `full`, when given, accumulates the ranges' lengths before pruning.
template <typename F>
__device__ __forceinline__ void for_each_range_row(const CellTable& T, int3 c, F&& fn, uint32_t* full = nullptr) {
    const uint32_t x0 = (uint32_t)c.x - 1u;
    const uint32_t rx = x0 - T.lo[0];
    const bool xin = T.dim[0] >= 3u && rx <= T.dim[0] - 3u;
    uint32_t hx[3], lx[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        hx[d] = (x0 + (uint32_t)d) * 19349663u;
        lx[d] = (x0 + (uint32_t)d) & 63u;
    }
#pragma unroll 1
    for (int row = 0; row < 9; ++row) {
        const uint32_t y = (uint32_t)c.y + (uint32_t)(row % 3 - 1);
        const uint32_t z = (uint32_t)c.z + (uint32_t)(row / 3 - 1);
        const uint32_t ry = y - T.lo[1], rz = z - T.lo[2];
        uint2 r[3];
        if (xin && ry < T.dim[1] && rz < T.dim[2]) {
            const uint2* q = T.grid + (rx + T.dim[0] * (ry + T.dim[1] * rz));
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                r[d] = q[d];
                const uint32_t far = r[d].x & kFarBucket;
                r[d].x ^= far;
                r[d].y &= ~kQuietCell;
                if (full) *full += r[d].y - r[d].x;
                if (far && T.prune) r[d].y = r[d].x;
            }
        } else {
            const uint32_t hyz = (y * 73856093u) ^ (z * 83492791u);
            const uint32_t lyz = 4u * (y & 63u) + 16u * (z & 63u);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                r[d] = T.cells[bucket_mod((((hx[d] ^ hyz) << 6) + lx[d] + lyz) & 0x7FFFFFFFu, T.g)];
                if (full) *full += r[d].y - r[d].x;
            }
        }
        fn(r[0], r[1], r[2]);
    }
}

// Whether cell c is inside T's box and marked kQuietCell.
__device__ __forceinline__ bool quiet_cell(const CellTable& T, int3 c) {
    const uint32_t rx = (uint32_t)c.x - T.lo[0], ry = (uint32_t)c.y - T.lo[1], rz = (uint32_t)c.z - T.lo[2];
    if (rx >= T.dim[0] || ry >= T.dim[1] || rz >= T.dim[2]) return false;
    return (T.grid[rx + T.dim[0] * (ry + T.dim[1] * rz)].y & kQuietCell) != 0;
}

template <typename F>
__device__ __forceinline__ void for_each_bucket(int3 c, const HashGrid& g, F&& fn) {
    for_each_bucket_row(c, g, [&](uint32_t k0, uint32_t k1, uint32_t k2) {
        fn(k0);
        fn(k1);
        fn(k2);
    });
}

}  // namespace pcisph
