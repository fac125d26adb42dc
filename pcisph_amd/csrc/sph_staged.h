// sph_staged.h - the fluid pressure pass with its candidates staged in LDS.
//
// The pressure pass (update_pressure, PCISPH/data/kernels/pcisph.cl:212-263) runs in
// every correction iteration and gathers, per fluid particle, the predicted position
// of every particle in the 27 buckets around its cell (~200 candidates at h = 2
// spacings): one divergent 16-byte load per candidate per lane.  Here the bookkeeping
// that lets a workgroup share those loads is done once per step, and each iteration
// only copies and walks:
//
//  * work order: fluid particles are taken in the Morton order of their own cell
//    (a stable sort of the slots by that code), so a workgroup's particles occupy a
//    compact block of cells even where a bucket holds particles of two far cells (hash
//    collisions place them with their bucket, not with their cell);
//  * staging map (per workgroup): for each (y, z) line of its box of cells, the run of
//    x-adjacent cells its particles' rows on that line cover; the cells' bucket ranges
//    (the reference's buckets, foreign particles included) are staged one after
//    another, so the three buckets of any particle's row are consecutive in LDS;
//  * row descriptors (per particle): LDS start and length of its nine (dz, dy) rows.
//
// Each iteration a workgroup copies its staged predicted positions into LDS (coalesced
// within each bucket) and every lane walks its nine rows there: the reference's
// candidates in the reference's order (rows dz-major then dy, x fastest within a row,
// ascending index within a bucket), so sums are unchanged bit for bit.  The boundary
// candidates of the same cells are staged after the fluid ones, as a pruning walk
// reads them (buckets flagged far left out): a particle whose walk may prune
// (CellTable) takes its nine boundary rows from LDS, the others walk global memory
// as in k_pressure; far buckets add exactly zero for the former, so both give the
// walk's sum.  The boundary ranges of
// the same launch (boundary mode) take their particles in the Morton order of their
// own cells (the boundary work order, sorted when the boundary is).  A workgroup whose
// staging does not fit (kStageCap candidates, kStageMaxCells cells, kStageLines lines,
// cell coordinates of magnitude 2^29 or more) walks global memory instead; one whose
// candidates exceed the staging cap (StageMaps::cap, at most kStageCap) but not twice
// it stages the lines that fit, and each particle walks a row lying wholly in LDS
// there and any other row's three buckets in global memory, in the same order.
#pragma once

#include "sph_kernels.h"

namespace pcisph {

constexpr uint32_t kRangeTileShift = 8;      // fluid ranges: their workgroup's staging index above the flag bits
constexpr uint32_t kStageCap = 2048;         // staged candidates per workgroup at most (x, y, z: 24 KB of LDS)
constexpr uint32_t kStageMaxCells = 2048;    // staged cells per workgroup
constexpr uint32_t kStageLines = 1024;       // (y, z) lines of a workgroup's box
constexpr uint32_t kStageFallback = 0xFFFFFFFFu;
constexpr int32_t kStageCoordLimit = 1 << 29;
constexpr int kCopyBatch = 8;   // staging loads in flight per thread
constexpr int kRangeBatch = 4;  // cell-range lookups in flight per thread (setup)

struct StageMaps {
    const uint32_t* work;  // [n_total] work index -> fluid slot (per-step Morton order of own cells)
    uint32_t* rows;        // [9][n_total] per work index: row r's LDS start | length << 16
    uint32_t* src;         // [workgroups][kStageCap] slot of each staged candidate (fluid, then boundary)
    uint32_t* info;        // [workgroups] fluid candidate count (the first kStageCap staged), or kStageFallback
    uint32_t* brows;       // [9][n_total] per work index: boundary row r's LDS start | length << 16
    uint32_t* binfo;       // [workgroups] staged boundary candidate count, or kStageFallback
    const uint32_t* bwork; // [nb_total] boundary work index -> boundary slot (Morton order of own cells)
    uint32_t cap;          // candidates staged per workgroup, 64..kStageCap; up to 2 cap partly staged
};

// The bucket range of cell (x, y, z) as for_each_range_row reads it (grid flags
// removed); with `prune`, a grid entry flagged kFarBucket reads as empty, as a pruning
// walk reads it.
__device__ __forceinline__ uint2 cell_range(const CellTable& T, uint32_t x, uint32_t y, uint32_t z,
                                            bool prune = false) {
    const uint32_t rx = x - T.lo[0], ry = y - T.lo[1], rz = z - T.lo[2];
    if (rx < T.dim[0] && ry < T.dim[1] && rz < T.dim[2]) {
        uint2 r = T.grid[rx + T.dim[0] * (ry + T.dim[1] * rz)];
        const uint32_t far = r.x & kFarBucket;
        r.x ^= far;
        r.y &= ~kQuietCell;
        if (far && prune) r.y = r.x;
        return r;
    }
    return T.cells[hash_key(x, y, z, T.g)];
}

// Exclusive workgroup scan of two counters (any multiple of 64 threads up to
// kMaxBlockThreads); `total` gets the sums.
__device__ __forceinline__ uint2 block_scan2(uint32_t a, uint32_t b, uint2& total, uint32_t (*part)[kMaxBlockThreads / 64]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = (int)(blockDim.x >> 6);
    uint32_t xa = a, xb = b;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t ya = __shfl_up(xa, off), yb = __shfl_up(xb, off);
        if (lane >= off) xa += ya, xb += yb;
    }
    if (lane == 63) part[0][wave] = xa, part[1][wave] = xb;
    __syncthreads();
    uint32_t ba = 0, bb = 0, ta = 0, tb = 0;
    for (int w = 0; w < waves; ++w) {
        if (w < wave) ba += part[0][w], bb += part[1][w];
        ta += part[0][w], tb += part[1][w];
    }
    __syncthreads();  // part may be reused by the caller
    total = make_uint2(ta, tb);
    return make_uint2(ba + xa - a, bb + xb - b);
}

// Work-order keys: the order code (order_code) of each particle's own cell.
template <typename R>
__global__ void k_stage_codes(const SphParams<R>* __restrict__ scenes, const BlockRange* __restrict__ blocks,
                              const V4<R>* __restrict__ pos, uint32_t bits, uint32_t* __restrict__ keys,
                              uint32_t* __restrict__ vals) {
    const BlockRange rg = blocks[blockIdx.x];
    const uint32_t i = rg.begin + threadIdx.x;
    if (i >= rg.end) return;
    const SphParams<R>& P = scenes[rg.scene];
    keys[i] = order_code<R>(P, rg.scene, cell_of4<R>(pos[i], P.cell), bits);
    vals[i] = i;
}

// Per step, one workgroup per fluid range of the block table (work indices
// [begin, end)): the staging map and the row descriptors of its particles, from the
// positions of the step (its buckets) and the fluid cell table.  Lines are the (y, z)
// rows of cells of the workgroup's bounding box of particle cells grown by one cell;
// each line stages the run of x-adjacent cells from the lowest x any particle's row
// on that line needs to the highest, so each particle's row of three cells is one run
// of LDS.
template <typename R>
__global__ __launch_bounds__(kMaxBlockThreads) void k_stage_setup(Batch<R> B, const uint2* __restrict__ fcells,
                                                                  const uint2* __restrict__ bcells,
                                                                  const V4<R>* __restrict__ pos, StageMaps S) {
    __shared__ int s_lx[kStageLines], s_hx[kStageLines];  // x extent of each line
    // 16-bit where the values that are read stay below 2^16 (cells <= kStageMaxCells,
    // staged offsets <= 2 cap); counts are clamped to 0xFFFF, which still sends the
    // workgroup to the fallback.  34 KB of LDS in all: four workgroups per CU.
    __shared__ uint16_t s_cell0[kStageLines + 1];         // first staged cell of each line
    __shared__ uint32_t s_start[kStageMaxCells];          // bucket start of each staged cell
    __shared__ uint16_t s_off[kStageMaxCells + 1];        // staged offset of each cell (count before the scan)
    __shared__ uint32_t s_bstart[kStageMaxCells];         // the same for the boundary buckets
    __shared__ uint16_t s_boff[kStageMaxCells + 1];
    __shared__ uint32_t s_part[2][kMaxBlockThreads / 64];
    __shared__ int s_box[6][kMaxBlockThreads / 64];
    const BlockRange rg = B.blocks[blockIdx.x];
    const uint32_t tile = rg.flags >> kRangeTileShift;
    const SphParams<R>& P = B.scenes[rg.scene];
    const CellTable FT = cell_table(P, fcells, B.fgrid);
    const CellTable BT = cell_table(P, bcells, B.bgrid);
    const uint32_t k = rg.begin + threadIdx.x;
    const bool act = k < rg.end;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = (int)(blockDim.x >> 6);
    int3 c = make_int3(0, 0, 0);
    if (act) c = cell_of4<R>(pos[S.work[k]], P.cell);
    // 1. the box of the workgroup's cells
    int lo[3] = {act ? c.x : INT32_MAX, act ? c.y : INT32_MAX, act ? c.z : INT32_MAX};
    int hi[3] = {act ? c.x : INT32_MIN, act ? c.y : INT32_MIN, act ? c.z : INT32_MIN};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            lo[a] = min(lo[a], __shfl_xor(lo[a], off));
            hi[a] = max(hi[a], __shfl_xor(hi[a], off));
        }
        if (lane == 0) {
            s_box[a][wave] = lo[a];
            s_box[3 + a][wave] = hi[a];
        }
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 3; ++a)
        for (int w = 0; w < waves; ++w) {
            lo[a] = min(lo[a], s_box[a][w]);
            hi[a] = max(hi[a], s_box[3 + a][w]);
        }
    bool fits = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) fits = fits && lo[a] > -kStageCoordLimit && hi[a] < kStageCoordLimit;
    const uint32_t by = fits ? (uint32_t)(hi[1] - lo[1] + 3) : 0u, bz = fits ? (uint32_t)(hi[2] - lo[2] + 3) : 0u;
    const uint64_t lines = (uint64_t)by * bz;
    if (!fits || lines > kStageLines) {
        if (threadIdx.x == 0) S.info[tile] = S.binfo[tile] = kStageFallback;
        return;
    }
    const uint32_t nl = (uint32_t)lines;
    const int y0 = lo[1] - 1, z0 = lo[2] - 1;
    for (uint32_t l = threadIdx.x; l < nl; l += blockDim.x) {
        s_lx[l] = INT32_MAX;
        s_hx[l] = INT32_MIN;
    }
    __syncthreads();
    // 2. x extent of each line: the first lane of each run of one cell (a cell's
    // particles are adjacent in the work order) records the rows of all of them
    const int px = __shfl_up(c.x, 1), py = __shfl_up(c.y, 1), pz = __shfl_up(c.z, 1);
    if (act && (lane == 0 || px != c.x || py != c.y || pz != c.z)) {
        for (int row = 0; row < 9; ++row) {
            const uint32_t l = (uint32_t)(c.z + row / 3 - 1 - z0) * by + (uint32_t)(c.y + row % 3 - 1 - y0);
            atomicMin(&s_lx[l], c.x - 1);
            atomicMax(&s_hx[l], c.x + 1);
        }
    }
    __syncthreads();
    // 3. first staged cell of each line: a contiguous chunk of lines per thread, scanned
    {
        const uint32_t chunk = (nl + blockDim.x - 1) / blockDim.x;
        const uint32_t l0 = min(threadIdx.x * chunk, nl), l1 = min(l0 + chunk, nl);
        uint32_t sum = 0;
        // a line of more than kStageMaxCells cells counts as kStageMaxCells + 1 (it sends
        // the workgroup to the fallback either way), so the sums cannot wrap
        auto cells_of = [&](uint32_t l) {
            return s_hx[l] >= s_lx[l] ? min((uint32_t)(s_hx[l] - s_lx[l]) + 1u, kStageMaxCells + 1u) : 0u;
        };
        for (uint32_t l = l0; l < l1; ++l) sum += cells_of(l);
        uint2 tot;
        const uint2 base = block_scan2(sum, 0u, tot, s_part);
        uint32_t run = base.x;
        for (uint32_t l = l0; l < l1; ++l) {
            s_cell0[l] = (uint16_t)run;
            run += cells_of(l);
        }
        if (threadIdx.x == 0) s_cell0[nl] = (uint16_t)min(tot.x, 0xFFFFu);
        if (tot.x > kStageMaxCells) {
            if (threadIdx.x == 0) S.info[tile] = S.binfo[tile] = kStageFallback;
            return;
        }
    }
    __syncthreads();
    const uint32_t ncells = s_cell0[nl];
    // 4. each staged cell's fluid bucket range and its boundary bucket range as a
    // pruning walk reads it (kRangeBatch lookups of each in flight per thread)
    const bool walls = P.nb != 0;
    for (uint32_t q0 = threadIdx.x; q0 < ncells; q0 += kRangeBatch * blockDim.x) {
        uint2 r[kRangeBatch], rb[kRangeBatch];
#pragma unroll
        for (int u = 0; u < kRangeBatch; ++u) {
            const uint32_t q = q0 + u * blockDim.x;
            r[u] = rb[u] = make_uint2(0u, 0u);
            if (q >= ncells) continue;
            uint32_t a = 0, b = nl;  // the last line whose first cell is <= q (a non-empty one)
            while (b - a > 1) {
                const uint32_t mid = (a + b) >> 1;
                if (s_cell0[mid] <= q) a = mid;
                else b = mid;
            }
            const uint32_t x = (uint32_t)(s_lx[a] + (int)(q - s_cell0[a]));
            const uint32_t y = (uint32_t)(y0 + (int)(a % by)), z = (uint32_t)(z0 + (int)(a / by));
            r[u] = cell_range(FT, x, y, z);
            if (walls) rb[u] = cell_range(BT, x, y, z, true);
        }
#pragma unroll
        for (int u = 0; u < kRangeBatch; ++u) {
            const uint32_t q = q0 + u * blockDim.x;
            if (q >= ncells) continue;
            s_start[q] = r[u].x;
            s_off[q] = (uint16_t)min(r[u].y - r[u].x, 0xFFFFu);
            s_bstart[q] = rb[u].x;
            s_boff[q] = (uint16_t)min(rb[u].y - rb[u].x, 0xFFFFu);
        }
    }
    __syncthreads();
    // 5. staged offsets of both: a contiguous chunk of cells per thread, scanned
    const uint32_t chunk = (ncells + blockDim.x - 1) / blockDim.x;
    const uint32_t q0 = min(threadIdx.x * chunk, ncells), q1 = min(q0 + chunk, ncells);
    uint32_t sum = 0, bsum = 0;
    for (uint32_t q = q0; q < q1; ++q) sum += s_off[q], bsum += s_boff[q];
    uint2 tot;
    const uint2 base = block_scan2(sum, bsum, tot, s_part);
    {
        uint32_t run = base.x, brun = base.y;
        for (uint32_t q = q0; q < q1; ++q) {
            const uint32_t cnt = s_off[q], bcnt = s_boff[q];
            s_off[q] = (uint16_t)run;
            s_boff[q] = (uint16_t)brun;
            run += cnt;
            brun += bcnt;
        }
    }
    const uint32_t total = tot.x;
    // the boundary candidates follow the fluid ones in the staging, if both fit
    const uint32_t btotal = total + tot.y <= S.cap ? tot.y : kStageFallback;
    if (threadIdx.x == 0) {
        s_off[ncells] = (uint16_t)total;
        s_boff[ncells] = (uint16_t)tot.y;
    }
    if (total > 2u * S.cap) {
        if (threadIdx.x == 0) S.info[tile] = S.binfo[tile] = kStageFallback;
        return;
    }
    __syncthreads();
    // 6. the staging map and the rows
    uint32_t* src = S.src + (size_t)tile * kStageCap;
    for (uint32_t q = threadIdx.x; q < ncells; q += blockDim.x) {
        const uint32_t o = s_off[q], n = s_off[q + 1] - o, b = s_start[q];
        const uint32_t fit = o < S.cap ? min(n, S.cap - o) : 0u;  // past the cap: not staged
        for (uint32_t e = 0; e < fit; ++e) src[o + e] = b + e;
        if (btotal != kStageFallback) {
            const uint32_t bo = total + s_boff[q], bn = s_boff[q + 1] - s_boff[q], bb = s_bstart[q];
            for (uint32_t e = 0; e < bn; ++e) src[bo + e] = bb + e;
        }
    }
    if (act) {
        for (int row = 0; row < 9; ++row) {
            const uint32_t l = (uint32_t)(c.z + row / 3 - 1 - z0) * by + (uint32_t)(c.y + row % 3 - 1 - y0);
            const uint32_t q = s_cell0[l] + (uint32_t)(c.x - 1 - s_lx[l]);
            const uint32_t a = s_off[q], e = s_off[q + 3];
            S.rows[(size_t)row * B.n_total + k] = a | ((e - a) << 16);
            if (btotal != kStageFallback && btotal) {
                const uint32_t ba = total + s_boff[q], be = total + s_boff[q + 3];
                S.brows[(size_t)row * B.n_total + k] = ba | ((be - ba) << 16);
            }
        }
    }
    if (threadIdx.x == 0) {
        S.info[tile] = total;
        S.binfo[tile] = btotal;
    }
}

// The poly6 terms of one staged row, LDS candidates [a, a + n) in order (x and y in
// `xy`, z in `z`), added to `pd`.  Four candidates per step from two LDS pointers
// (paired loads, constant offsets), each step one add per pointer and one compare
// against the row's last full step.
template <bool FAST>
__device__ __forceinline__ float staged_row(const float2* __restrict__ xy, const float* __restrict__ z, uint32_t a,
                                            uint32_t n, const float4& pr, float h2, float pd) {
    const float2* q = xy + a;
    const float* qz = z + a;
    const float2* const end4 = q + (n & ~3u);
    const float2* const end = q + n;
    for (; q != end4; q += 4, qz += 4) {
        float4 o[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) o[u] = make_float4(q[u].x, q[u].y, qz[u], 0.f);
#pragma unroll
        for (int u = 0; u < 4; ++u) pd = poly6_add<float, FAST>(pd, pr, o[u], h2);
    }
    for (; q != end; ++q, ++qz) pd = poly6_add<float, FAST>(pd, pr, make_float4(q->x, q->y, *qz, 0.f), h2);
    return pd;
}

// The same row of a partly staged workgroup that lies past the staging: row `row` of
// cell c's 27 (dz-major, then dy) with its three bucket ranges read from the cell
// table and their candidates' predicted positions from global memory.
template <bool FAST>
__device__ __forceinline__ float global_row(const CellTable& T, int3 c, int row, const float4* __restrict__ xp,
                                            const float4& pr, float h2, float pd) {
    const uint32_t x = (uint32_t)c.x - 1u, y = (uint32_t)c.y + (uint32_t)(row % 3 - 1),
                   z = (uint32_t)c.z + (uint32_t)(row / 3 - 1);
    gather_row<kGatherBatch>(cell_range(T, x, y, z), cell_range(T, x + 1u, y, z), cell_range(T, x + 2u, y, z),
                             [&](uint32_t j) { return xp[j]; },
                             [&](const float4& o, uint32_t) { pd = poly6_add<float, FAST>(pd, pr, o, h2); });
    return pd;
}

// The candidates of one staged row as a walk visits them: LDS entries [a, a + n) of
// `s` (x, y, z and the candidate's slot in w), use(o, slot) in order, four at a time.
template <typename Use>
__device__ __forceinline__ void staged_row_each(const float4* __restrict__ s, uint32_t a, uint32_t n, Use&& use) {
    const float4* q = s + a;
    const float4* const end4 = q + (n & ~3u);
    const float4* const end = q + n;
    for (; q != end4; q += 4) {
        float4 o[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) o[u] = q[u];
#pragma unroll
        for (int u = 0; u < 4; ++u) use(o[u], __float_as_uint(o[u].w));
    }
    for (; q != end; ++q) {
        const float4 o = *q;
        use(o, __float_as_uint(o.w));
    }
}

// Row `row` of cell c's 27 (dz-major, then dy) walked in global memory: its three
// bucket ranges from the cell table, candidates through load(j), use(data, j) in order.
template <typename Load, typename Use>
__device__ __forceinline__ void global_row_each(const CellTable& T, int3 c, int row, Load&& load, Use&& use) {
    const uint32_t x = (uint32_t)c.x - 1u, y = (uint32_t)c.y + (uint32_t)(row % 3 - 1),
                   z = (uint32_t)c.z + (uint32_t)(row / 3 - 1);
    gather_row<kGatherBatch>(cell_range(T, x, y, z), cell_range(T, x + 1u, y, z), cell_range(T, x + 2u, y, z), load,
                             use);
}

// update_density (pcisph.cl:65-92) of the FP32 step, with the same counts and neighbour
// lists as k_density, its fluid ranges walking the step's LDS staging (see above) at
// the positions x instead of the 27 buckets in global memory: the workgroup copies its
// staged candidates' positions and slots into LDS (x, y, z, slot), then each particle
// of its work order walks its nine rows there -- the reference's candidates in the
// reference's order, so the sums, counts and lists are the global walk's bit for bit.
// Its boundary walk reads the staged boundary rows where a pruning walk may (the
// candidate count still takes the pruned buckets, as walk_buckets' `full`), global
// memory otherwise; a partly staged workgroup walks the rows past the staging in
// global memory, a workgroup without staging all of them.  Boundary ranges as
// k_density.
template <bool FAST>
__global__ __launch_bounds__(kMaxBlockThreads) void k_density_staged(
    Batch<float> B, const uint2* __restrict__ bcells, const float4* __restrict__ bpos,
    const uint2* __restrict__ fcells, const float4* __restrict__ pos, float4* __restrict__ xr,
    uint4* __restrict__ counts, float* __restrict__ bp, float4* __restrict__ bxf, NeighborLists NL, StageMaps S) {
    __shared__ float4 s_x[kStageCap];
    __shared__ unsigned long long lds[16];
    const BlockRange rg = B.blocks[blockIdx.x];
    const SphParams<float> P = B.scenes[rg.scene];
    LoopState* st = B.states + rg.scene;
    const CellTable FT = cell_table(P, fcells, B.fgrid);
    const bool lists = NL.on != 0;
    const float lr2 = P.list_r2;
    if (lists) clear_table(NL.disp[0], NL.buckets);  // filled by the force pass
    unsigned long long cf_sum = 0, cb_sum = 0;
    if (rg.flags & kRangeBoundary) {
        const uint32_t t = rg.begin + threadIdx.x;
        if (t < rg.end) density_boundary_particle<float>(B, P, t, FT, bpos, pos, bp, bxf, NL);
    } else {
        const uint32_t tile = rg.flags >> kRangeTileShift;
        const uint32_t total = S.info[tile];
        const bool staged = total != kStageFallback;
        const uint32_t btotal = staged ? S.binfo[tile] : kStageFallback;
        const bool bstaged = btotal != kStageFallback;
        const uint32_t fitted = staged ? min(total, S.cap) : 0u;
        if (staged) {
            const uint32_t all = fitted + (bstaged ? btotal : 0u);
            const uint32_t* src = S.src + (size_t)tile * kStageCap;
            for (uint32_t q0 = threadIdx.x; q0 < all; q0 += kCopyBatch * blockDim.x) {
                uint32_t j[kCopyBatch];
                float4 v[kCopyBatch];
                const float4* from[kCopyBatch];
#pragma unroll
                for (int u = 0; u < kCopyBatch; ++u) {
                    const uint32_t q = q0 + u * blockDim.x, qq = q < all ? q : q0;
                    j[u] = src[qq];
                    from[u] = qq < fitted ? pos : bpos;
                }
#pragma unroll
                for (int u = 0; u < kCopyBatch; ++u) v[u] = from[u][j[u]];
#pragma unroll
                for (int u = 0; u < kCopyBatch; ++u) {
                    const uint32_t q = q0 + u * blockDim.x;
                    if (q < all) s_x[q] = make_float4(v[u].x, v[u].y, v[u].z, __uint_as_float(j[u]));
                }
            }
        }
        __syncthreads();
        const uint32_t k = rg.begin + threadIdx.x;
        if (k < rg.end) {
            const uint32_t i = S.work[k];
            const float4 self = pos[i];
            const int3 c = cell_of4<float>(self, P.cell);
            float rho = 0.f;
            uint32_t cf = 0, cb = 0, nf = 0, nb = 0;
            ListWriter wb{NL.b, B.n_total, i, NL.cap_b, 0, {}}, wf{NL.f, B.n_total, i, NL.cap_f, 0, {}};
            // the terms and counts of k_density, candidate by candidate
            auto bnd = [&](const float4& o, uint32_t j) {
                float dx, dy, dz;
                const float r2 = dist2<float>(self.x, self.y, self.z, o.x, o.y, o.z, dx, dy, dz);
                rho += poly6<float>(r2, P.h2);
                nb += !(r2 > P.h2);
                if (lists && r2 <= lr2) wb.push(j);
            };
            auto fl = [&](const float4& o, uint32_t j) {
                float dx, dy, dz;
                const float r2 = dist2<float>(self.x, self.y, self.z, o.x, o.y, o.z, dx, dy, dz);
                rho += poly6<float>(r2, P.h2);
                ++cf;
                nf += !(r2 > P.h2);
                if (lists && r2 <= lr2) wf.push(j);
            };
            auto bload = [&](uint32_t j) { return bpos[j]; };
            auto fload = [&](uint32_t j) { return pos[j]; };
            if (P.nb) {
                // terms vanish beyond h; list entries beyond the list radius
                CellTable BT = cell_table(P, bcells, B.bgrid);
                BT.prune = far_prunable<float>(0.f, sqrtf(fmaxf((float)P.h2, lists ? lr2 : 0.f)), P.cell);
                if (BT.prune && bstaged) {
                    for_each_range_row(BT, c, [](uint2, uint2, uint2) {}, &cb);
                    if (btotal) {
#pragma unroll 1
                        for (int row = 0; row < 9; ++row) {
                            const uint32_t d = S.brows[(size_t)row * B.n_total + k];
                            staged_row_each(s_x, d & 0xFFFFu, d >> 16, bnd);
                        }
                    }
                } else {
                    walk_buckets<float>(BT, c, bload, bnd, &cb);
                }
            }
            if (staged) {
#pragma unroll 1
                for (int row = 0; row < 9; ++row) {
                    const uint32_t d = S.rows[(size_t)row * B.n_total + k];
                    const uint32_t a = d & 0xFFFFu, n = d >> 16;
                    if (a + n <= fitted) staged_row_each(s_x, a, n, fl);
                    else global_row_each(FT, c, row, fload, fl);
                }
            } else {
                walk_buckets<float>(FT, c, fload, fl);
            }
            rho *= P.density_scale;
            float4 out = self;
            out.w = rho;
            xr[i] = out;
            if (counts) counts[i] = make_uint4(cf, cb, nf, nb);
            if (lists) {
                NL.bcount[i] = P.nb ? wb.finish() : 0u;
                NL.fcount[i] = wf.finish();
            }
            cf_sum = cf;
            cb_sum = cb;
        }
    }
    const unsigned long long a = block_reduce_sum(cf_sum, lds);
    __syncthreads();
    const unsigned long long b = block_reduce_sum(cb_sum, lds + 8);
    if (threadIdx.x == 0 && (a || b)) {
        atomicAdd(&st->fluid_candidates, a);
        atomicAdd(&st->boundary_candidates, b);
    }
}

// The poly6 terms of one staged row for both starts of the dual pass: candidates
// [a, a + n) of the network start's staging (xyw, zw) against pw and of the plain
// start's (xy0, z0) against p0, added to s.x and s.y.
template <bool FAST>
__device__ __forceinline__ float2 staged_row_dual(const float2* __restrict__ xyw, const float* __restrict__ zw,
                                                  const float2* __restrict__ xy0, const float* __restrict__ z0,
                                                  uint32_t a, uint32_t n, const float4& pw, const float4& p0,
                                                  float h2, float2 s) {
    const float2* const end4 = xyw + a + (n & ~3u);
    const float2* const end = xyw + a + n;
    const float2 *q = xyw + a, *r = xy0 + a;
    const float *qz = zw + a, *rz = z0 + a;
    for (; q != end4; q += 4, r += 4, qz += 4, rz += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            s.x = poly6_add<float, FAST>(s.x, pw, make_float4(q[u].x, q[u].y, qz[u], 0.f), h2);
            s.y = poly6_add<float, FAST>(s.y, p0, make_float4(r[u].x, r[u].y, rz[u], 0.f), h2);
        }
    }
    for (; q != end; ++q, ++r, ++qz, ++rz) {
        s.x = poly6_add<float, FAST>(s.x, pw, make_float4(q->x, q->y, *qz, 0.f), h2);
        s.y = poly6_add<float, FAST>(s.y, p0, make_float4(r->x, r->y, *rz, 0.f), h2);
    }
    return s;
}

// The dual first iteration of a warm-started step (k_pressure_dual: both starts'
// pressure passes in one) for the FP32 step, its fluid ranges from the step's LDS
// staging: both predictions of every staged fluid candidate in LDS (48 KB at the
// staging cap), the boundary candidates after them, each lane's nine rows walked once
// for both sums in the reference's order -- the global walk's sums bit for bit.
// Partly staged and unstaged workgroups walk global memory as k_pressure_dual does;
// boundary ranges as k_pressure_dual.
template <bool FAST>
__global__ __launch_bounds__(kMaxBlockThreads) void k_pressure_dual_staged(
    Batch<float> B, const uint2* __restrict__ bcells, const float4* __restrict__ bpos, const float* __restrict__ binit,
    const uint2* __restrict__ fcells, const float4* __restrict__ xr, const float4* __restrict__ xp,
    float* __restrict__ dv, float* __restrict__ p, float4* __restrict__ xf, float* __restrict__ bp,
    float4* __restrict__ bxf, DualArrays<float> D, StageMaps S) {
    __shared__ float2 s_xyw[kStageCap];
    __shared__ float s_zw[kStageCap];
    __shared__ float2 s_xy0[kStageCap];
    __shared__ float s_z0[kStageCap];
    __shared__ unsigned long long lds[4][8];
    if (D.disp_next) clear_table(D.disp_next, D.buckets);  // for iteration 1's prediction
    const BlockRange rg = B.blocks[blockIdx.x];
    const SphParams<float> P = B.scenes[rg.scene];
    if (!P.warm) return;  // scenes left out of the warm start: the plain pass runs iteration 0
    LoopState* st = B.states + rg.scene;
    const CellTable FT = cell_table(P, fcells, B.fgrid);
    unsigned long long m[4] = {0, 0, 0, 0};  // dv warm, exp warm, dv plain, exp plain
    struct Pair {
        float4 w, z;
    };
    if (rg.flags & kRangeBoundary) {
        const uint32_t t = rg.begin + threadIdx.x;
        if (t < rg.end) dual_boundary_particle<float, FAST>(P, t, FT, bpos, binit, xp, bp, bxf, D);
    } else {
        const uint32_t tile = rg.flags >> kRangeTileShift;
        const uint32_t total = S.info[tile];
        const bool staged = total != kStageFallback;
        const uint32_t btotal = staged ? S.binfo[tile] : kStageFallback;
        const bool bstaged = btotal != kStageFallback;
        const uint32_t fitted = staged ? min(total, S.cap) : 0u;
        if (staged) {
            const uint32_t all = fitted + (bstaged ? btotal : 0u);
            const uint32_t* src = S.src + (size_t)tile * kStageCap;
            for (uint32_t q0 = threadIdx.x; q0 < all; q0 += kCopyBatch * blockDim.x) {
                uint32_t j[kCopyBatch];
                float4 v[kCopyBatch], v0[kCopyBatch];
#pragma unroll
                for (int u = 0; u < kCopyBatch; ++u) {
                    const uint32_t q = q0 + u * blockDim.x;
                    j[u] = src[q < all ? q : q0];
                }
#pragma unroll
                for (int u = 0; u < kCopyBatch; ++u) {
                    const uint32_t q = q0 + u * blockDim.x, qq = q < all ? q : q0;
                    if (qq < fitted) {
                        v[u] = xp[j[u]];
                        v0[u] = D.xp0[j[u]];
                    } else {
                        v[u] = v0[u] = bpos[j[u]];
                    }
                }
#pragma unroll
                for (int u = 0; u < kCopyBatch; ++u) {
                    const uint32_t q = q0 + u * blockDim.x;
                    if (q < all) {
                        s_xyw[q] = make_float2(v[u].x, v[u].y);
                        s_zw[q] = v[u].z;
                        s_xy0[q] = make_float2(v0[u].x, v0[u].y);
                        s_z0[q] = v0[u].z;
                    }
                }
            }
        }
        __syncthreads();
        const uint32_t k = rg.begin + threadIdx.x;
        if (k < rg.end) {
            const uint32_t i = S.work[k];
            const float4 self = xr[i];
            const float4 pw = xp[i], pz = D.xp0[i];
            const int3 c = cell_of4<float>(self, P.cell);
            float2 sum = make_float2(0.f, 0.f);
            auto both = [&](const float4& o, uint32_t) {
                sum.x = poly6_add<float, FAST>(sum.x, pw, o, P.h2);
                sum.y = poly6_add<float, FAST>(sum.y, pz, o, P.h2);
            };
            auto pair = [&](const Pair& o, uint32_t) {
                sum.x = poly6_add<float, FAST>(sum.x, pw, o.w, P.h2);
                sum.y = poly6_add<float, FAST>(sum.y, pz, o.z, P.h2);
            };
            auto pload = [&](uint32_t j) { return Pair{xp[j], D.xp0[j]}; };
            if (P.nb) {
                CellTable BT = cell_table(P, bcells, B.bgrid);
                BT.prune = far_prunable<float>(displacement<float>(pw, self), (float)P.h, P.cell) &&
                           far_prunable<float>(displacement<float>(pz, self), (float)P.h, P.cell);
                if (BT.prune && quiet_cell(BT, c)) {
                    // nothing within reach
                } else if (BT.prune && bstaged) {
                    if (btotal) {
#pragma unroll 1
                        for (int row = 0; row < 9; ++row) {
                            const uint32_t d = S.brows[(size_t)row * B.n_total + k];
                            sum = staged_row_dual<FAST>(s_xyw, s_zw, s_xyw, s_zw, d & 0xFFFFu, d >> 16, pw, pz,
                                                        P.h2, sum);
                        }
                    }
                } else {
                    walk_buckets<float>(BT, c, [&](uint32_t j) { return bpos[j]; }, both);
                }
            }
            if (staged) {
#pragma unroll 1
                for (int row = 0; row < 9; ++row) {
                    const uint32_t d = S.rows[(size_t)row * B.n_total + k];
                    const uint32_t a = d & 0xFFFFu, n = d >> 16;
                    if (a + n <= fitted) sum = staged_row_dual<FAST>(s_xyw, s_zw, s_xy0, s_z0, a, n, pw, pz, P.h2, sum);
                    else global_row_each(FT, c, row, pload, pair);
                }
            } else {
                walk_buckets<float>(FT, c, pload, pair);
            }
            dual_fluid_outputs<float, FAST>(P, i, self, sum.x, sum.y, dv, p, xf, D, m);
        }
    }
    // 0: dv warm, 1: expansion warm, 2: dv plain, 3: expansion plain
    unsigned long long* const dst[4] = {&st->dual_dv[0], &st->dual_exp[0], &st->dual_dv[1], &st->dual_exp[1]};
    block_max_to<4>(m, dst, lds);
}

// update_pressure (pcisph.cl:212-263), the fluid ranges with their candidates from the
// workgroup's LDS staging (see above) -- the same results as k_pressure's bucket walk --
// and the boundary ranges (boundary mode) walking their fluid buckets in global
// memory, in one launch so both kinds of workgroups share the device.  Fluid `blocks`
// entries carry their staging index (kRangeTileShift); boundary ranges are taken in
// the boundary work order (S.bwork), so a wavefront's 27-cell neighbourhoods overlap
// instead of lying wherever the hash put their buckets.
template <bool FAST>
__device__ __forceinline__ void pressure_staged_range(
    const Batch<float>& B, uint32_t range, uint32_t iter, const uint2* __restrict__ bcells,
    const float4* __restrict__ bpos, const float* __restrict__ binit, const uint2* __restrict__ fcells,
    const float4* __restrict__ xr, const float4* __restrict__ xp, float* __restrict__ dv, float* __restrict__ p,
    float4* __restrict__ xf, float* __restrict__ bp, float4* __restrict__ bxf, const StageMaps& S) {
    using O = Ops<float, FAST>;
    __shared__ float2 s_xy[kStageCap];
    __shared__ float s_z[kStageCap];
    __shared__ unsigned long long lds[2][8];
    const BlockRange rg = B.blocks[range];
    const SphParams<float> P = B.scenes[rg.scene];
    LoopState* st = B.states + rg.scene;
    if (pressure_pass_skipped<float>(st, iter, P)) return;
    if (rg.flags & kRangeBoundary) {
        // boundary mode (pcisph.cl:231-236): binit + the fluid candidates at x^
        const uint32_t t = rg.begin + threadIdx.x;
        if (t < rg.end) {
            const uint32_t b = S.bwork[t];
            const float4 self = bpos[b];
            const int3 c = cell_of4<float>(self, P.cell);
            float pd = 0.f;
            pd += binit[b];
            walk_buckets<float>(cell_table(P, fcells, B.fgrid), c, [&](uint32_t j) { return xp[j]; },
                                [&](const float4& o, uint32_t) { pd = poly6_add<float, FAST>(pd, self, o, P.h2); });
            pd *= P.density_scale;
            const float v = fmaxf(0.f, pd - P.rho0);
            if (v != 0.f) {
                const float pb = bp[b] + v * P.delta;
                bp[b] = pb;
                float4 out = self;
                out.w = O::div(pb, P.rho0_sq);
                bxf[b] = out;
            }
        }
        return;  // boundary ranges take no part in the fluid maxima
    }
    const uint32_t tile = rg.flags >> kRangeTileShift;
    const uint32_t total = S.info[tile];
    const bool staged = total != kStageFallback;
    const uint32_t btotal = staged ? S.binfo[tile] : kStageFallback;
    const bool bstaged = btotal != kStageFallback;
    const uint32_t fitted = staged ? min(total, S.cap) : 0u;  // fluid candidates in LDS
    if (staged) {
        // kCopyBatch independent loads in flight per thread (indices, then positions):
        // the fluid candidates' predicted positions, then the boundary candidates'
        // (staged only when every fluid candidate is)
        const uint32_t all = fitted + (bstaged ? btotal : 0u);
        const uint32_t* src = S.src + (size_t)tile * kStageCap;
        for (uint32_t q0 = threadIdx.x; q0 < all; q0 += kCopyBatch * blockDim.x) {
            uint32_t j[kCopyBatch];
            float4 v[kCopyBatch];
            const float4* from[kCopyBatch];
#pragma unroll
            for (int u = 0; u < kCopyBatch; ++u) {
                const uint32_t q = q0 + u * blockDim.x, qq = q < all ? q : q0;
                j[u] = src[qq];
                from[u] = qq < fitted ? xp : bpos;
            }
#pragma unroll
            for (int u = 0; u < kCopyBatch; ++u) v[u] = from[u][j[u]];
#pragma unroll
            for (int u = 0; u < kCopyBatch; ++u) {
                const uint32_t q = q0 + u * blockDim.x;
                if (q < all) {
                    s_xy[q] = make_float2(v[u].x, v[u].y);
                    s_z[q] = v[u].z;
                }
            }
        }
    }
    __syncthreads();
    const uint32_t k = rg.begin + threadIdx.x;
    unsigned long long m[2] = {0, 0};  // max dv, max expansion
    if (k < rg.end) {
        const uint32_t i = S.work[k];
        const float4 self = xr[i], pr = xp[i];
        const int3 c = cell_of4<float>(self, P.cell);
        float pd = 0.f;
        auto add = [&](const float4& o, uint32_t) { pd = poly6_add<float, FAST>(pd, pr, o, P.h2); };
        if (P.nb) {
            CellTable BT = cell_table(P, bcells, B.bgrid);
            BT.prune = far_prunable<float>(displacement<float>(pr, self), (float)P.h, P.cell);
            if (BT.prune && quiet_cell(BT, c)) {
                // nothing within reach
            } else if (BT.prune && bstaged) {
                // the pruning walk's candidates, staged after the fluid ones (none: nothing to add)
                if (btotal) {
                    uint32_t d[9];
#pragma unroll
                    for (int row = 0; row < 9; ++row) d[row] = S.brows[(size_t)row * B.n_total + k];
#pragma unroll 1
                    for (int row = 0; row < 9; ++row)
                        pd = staged_row<FAST>(s_xy, s_z, d[row] & 0xFFFFu, d[row] >> 16, pr, P.h2, pd);
                }
            } else {
                walk_buckets<float>(BT, c, [&](uint32_t j) { return bpos[j]; }, add);
            }
        }
        if (staged) {
            uint32_t d[9];
#pragma unroll
            for (int row = 0; row < 9; ++row) d[row] = S.rows[(size_t)row * B.n_total + k];
            if (total <= S.cap) {
#pragma unroll 1
                for (int row = 0; row < 9; ++row)
                    pd = staged_row<FAST>(s_xy, s_z, d[row] & 0xFFFFu, d[row] >> 16, pr, P.h2, pd);
            } else {
                const CellTable FT = cell_table(P, fcells, B.fgrid);
#pragma unroll 1
                for (int row = 0; row < 9; ++row) {
                    const uint32_t a = d[row] & 0xFFFFu, n = d[row] >> 16;
                    if (a + n <= S.cap) pd = staged_row<FAST>(s_xy, s_z, a, n, pr, P.h2, pd);
                    else pd = global_row<FAST>(FT, c, row, xp, pr, P.h2, pd);
                }
            }
        } else {
            walk_buckets<float>(cell_table(P, fcells, B.fgrid), c, [&](uint32_t j) { return xp[j]; }, add);
        }
        pd *= P.density_scale;
        const float v = fmaxf(0.f, pd - P.rho0);
        // expansion counts only where the step began at or above rest density
        if (self.w >= P.rho0) m[1] = Bits<float>::of(fmaxf(0.f, P.rho0 - pd));
        dv[i] = v;
        float pi = p[i];
        if (v != 0.f) {
            pi += v * P.delta;
            p[i] = pi;
        }
        float4 out = self;
        out.w = O::div(pi, self.w * self.w);
        xf[i] = out;
        m[0] = Bits<float>::of(v);
    }
    unsigned long long* const dst[2] = {&st->max_dv[iter], &st->max_exp[iter]};
    block_max_to<2>(m, dst, lds);
}

template <bool FAST>
__global__ __launch_bounds__(kMaxBlockThreads) void k_pressure_staged(
    Batch<float> B, uint32_t iter, const uint2* __restrict__ bcells, const float4* __restrict__ bpos,
    const float* __restrict__ binit, const uint2* __restrict__ fcells, const float4* __restrict__ xr,
    const float4* __restrict__ xp, float* __restrict__ dv, float* __restrict__ p, float4* __restrict__ xf,
    float* __restrict__ bp, float4* __restrict__ bxf, StageMaps S, NeighborLists NL) {
    if (NL.on) clear_table(NL.disp[(iter + 1) & 1], NL.buckets);  // for the next prediction
    pressure_staged_range<FAST>(B, blockIdx.x, iter, bcells, bpos, binit, fcells, xr, xp, dv, p, xf, bp, bxf, S);
}

}  // namespace pcisph
