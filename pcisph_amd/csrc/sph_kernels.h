// sph_kernels.h - the PCISPH neighbour and stream kernels for gfx950.
//
// Included by two translation units that differ only in floating-point contract:
//   sph_ieee.hip  (-ffp-contract=off): IEEE single/double, one rounding per
//                 operation, correctly rounded `/` and sqrt, the reference's order of
//                 operations and summation.  FP32 here reproduces the CPU oracle bit
//                 for bit; FP64 is the double-precision reference run.
//   sph_fast.hip  (-ffp-contract=fast): FP32 with fused multiply-adds and hardware
//                 reciprocal / square root, the accuracy class OpenCL allows the
//                 reference's `/`, distance() and fast_length().
//
// Layout in HBM (one array per field, structure of arrays, 16-byte elements so a
// candidate costs one dwordx4 load):
//   pos   float4 (x, y, z, 0)          positions of this step, in its storage order
//   xr    float4 (x, y, z, rho)        written by the density pass
//   vel, nrm, fo, fp, xp  float4       velocity, normal, other force, pressure force,
//                                      predicted position
//   xf    float4 (x, y, z, p/rho^2)    written by the pressure pass for the force pass
//   bpos  float4 (x, y, z, 0), bxf float4 (x, y, z, p_b/rho0^2) for the boundary
// Every kernel is one work-item per particle in storage order (below), so the 64
// lanes of a wavefront walk neighbouring buckets together and share their loads.
//
// Storage order.  The reference keeps its arrays in bucket order (sorted by cell
// hash), where consecutive buckets are cells scattered over the whole domain.  Here
// each bucket's fluid particles stay contiguous and in the reference's order, but
// the buckets follow a Morton curve of their cells (the Morton layout, built every
// step after the reference's sort), so the candidates a neighbourhood gathers lie
// together in memory and a wavefront covers a compact block of cells.  A walk over
// bucket k visits exactly the reference's candidates in the reference's order.  The
// solver maps storage slots to canonical indices (the reference's) at the public
// interface; with the layout off, storage order is the reference's.
//
// Batches.  A launch covers every scene of a batch.  The scenes' particles are
// concatenated (scene s owns fluid [f_off, f_off + n) and boundary
// [b_off, b_off + nb)), their bucket tables too (scene s owns buckets
// [bucket_base, bucket_base + B)), and a block table gives each workgroup one scene
// and one contiguous range of its fluid or boundary particles.  The scene's
// constants are therefore uniform across the workgroup and read with scalar loads,
// and a batch of one scene is the single-scene solver.  Sorting the concatenated
// keys bucket_base + hash keeps every scene contiguous and in the order its own
// sort would give (the sort is stable and the scenes' key ranges are disjoint and
// increasing).
#pragma once

#include "device_common.h"

namespace pcisph {

constexpr int kMaxIterationsCap = 256;

// Step-constant scalars of one scene, each product formed once on the host in the
// same order as the reference kernels write it (e.g. `density *= particle_mass *
// poly6_normalization`, PCISPH/data/kernels/pcisph.cl:90).
template <typename R>
struct SphParams {
    R h, h2, cell, rho0, m, dt, delta, threshold;
    R density_scale;   // m * K6                       (pcisph.cl:90, :252)
    R normal_scale;    // h * m * K6'                  (pcisph.cl:109)
    R visc_scale;      // mu * m * Kv                  (pcisph.cl:160)
    R coh_scale;       // -gamma * m * m * Kst          (pcisph.cl:162)
    R curv_scale;      // -gamma * m                    (pcisph.cl:163)
    R pf_scale;        // -Ks * m * m                   (pcisph.cl:297)
    R st_term;         // h^6 / 64
    R gravity_force;   // m * g                         (pcisph.cl:167)
    R rho0_sq;         // rho0 * rho0                   (pcisph.cl:282-283)
    HashGrid grid;
    uint32_t n, nb;                  // this scene's fluid / boundary counts
    uint32_t f_off, b_off;           // its first fluid / boundary particle in the batch
    uint32_t bucket_base;            // its first bucket in the batch's cell tables
    uint32_t min_iterations, max_iterations;
    float p_ref;                     // delta * rho0 / 100: warm-start pressure unit
    float div_scale;                 // m * 45 / (pi h^6) * dt * 100: warm-start divergence unit
    float list_r2;                   // neighbour-list radius squared, ((1 + skin) h)^2
    float disp_limit;                // 0.9 skin h: the largest d_i + D_i with which particle i may use its lists
    uint32_t box_lo[3];              // cell box of the dense range grids (CellTable); box_dim[0] == 0: none
    uint32_t box_dim[3];
    uint32_t box_base;               // its first entry in the batch's grids
    uint32_t warm;                   // 1: this step starts this scene from the network (a warm-started
                                     // step, the scene not left out of it); 0 in every other step
};

// The low `bits` bits of v spread to every third bit (bit k -> bit 3k).
__device__ __forceinline__ uint32_t spread_bits3(uint32_t v, uint32_t bits) {
    uint32_t out = 0;
    for (uint32_t k = 0; k < bits; ++k) out |= ((v >> k) & 1u) << (3 * k);
    return out;
}

// One workgroup's share of a launch: particles [begin, end) of one scene.
enum : uint32_t { kRangeBoundary = 1u, kRangeFirstOfScene = 2u };
struct BlockRange {
    uint32_t scene;
    uint32_t begin, end;  // indices into the batch's fluid (or boundary) arrays
    uint32_t flags;       // kRangeBoundary, kRangeFirstOfScene (first fluid range of its scene)
};

// Per-step device record.  max_dv[k] holds the bits of the non-negative maximum
// density variation of iteration k (integer order = float order for x >= 0).
struct LoopState {
    unsigned long long max_dv[kMaxIterationsCap];
    unsigned long long max_exp[kMaxIterationsCap];  // max(0, rho0 - rho*) over particles that began the
                                                    // step at or above rho0: the other side of the residual
    unsigned long long fluid_candidates;
    unsigned long long boundary_candidates;
    // iteration 0 of a warm-started step evaluates both starts: [0] warm, [1] plain
    unsigned long long dual_dv[2];
    unsigned long long dual_exp[2];
    unsigned int nonfinite;
    unsigned int warm_rejected;  // 1: the plain start had the smaller residual
    unsigned int list_waves;     // fluid wavefronts of the pressure passes that walked neighbour lists
    unsigned int pad;
};

// Second set of per-particle arrays for the dual first iteration of a warm start:
// the plain start's predicted positions, pressures, p/rho^2 and density variations.
template <typename R>
struct DualArrays {
    V4<R>* xp0;
    V4<R>* xf0;
    R* p0;
    R* dv0;
    R* bp0;
    V4<R>* bxf0;
    uint32_t* disp_next;  // the neighbour lists' displacement table of iteration 1 (or null)
    uint32_t buckets;
    uint32_t pad;
};

// Host-visible record of the correction loop (pinned, mapped memory), written by the
// first workgroup of each pressure-force launch: per iteration, how many scenes go on
// after it, then the step tag that marks the iteration's results as published.
struct HostMirror {
    unsigned int running[kMaxIterationsCap];
    unsigned int tag[kMaxIterationsCap];
    unsigned int sort_abort;  // raised by a one-sweep sort whose look-back wait expired
};

// Per-scene results, host-visible, rewritten at every iteration the scene runs, so
// after the loop they hold the scene's last iteration.
struct SceneResult {
    unsigned long long max_dv;   // bits of the max density variation (precision R)
    unsigned long long max_exp;  // bits of the max expansion
    unsigned long long candidates[2];
    unsigned long long first_dv;  // bits of the plain start's max density variation after iteration 0
    unsigned int iterations;
    unsigned int warm_rejected;
    unsigned int list_waves;
    unsigned int running;  // 1: the scene runs another iteration after this one
};

// Inputs of the warm-start network, per fluid particle (8 floats, row-major):
//  0 (rho/rho0 - 1) * 100          1 fluid neighbours / 32
//  2 previous-step pressure / p_ref 3 |v| dt / h * 100
//  4 neighbour mean of input 0      5 neighbour mean of input 2
//  6 boundary neighbours / 32       7 velocity divergence * dt * 100
// with p_ref = delta * rho0 / 100, the pressure one iteration assigns to a 1% error.
struct FeatureArgs {
    float* feat;          // float[n][8] or null (no features)
    const uint4* counts;  // per-particle neighbour counts of the density pass
    uint32_t all;         // 1: every scene's (samples are collected), 0: the warm-started scenes' only
};

// The batch as every step kernel sees it.  Neighbour lists and their counts are
// kept per storage slot (slot-major), so a wavefront reads its lists coalesced.
template <typename R>
struct Batch {
    const SphParams<R>* scenes;   // [S]
    const BlockRange* blocks;     // fluid ranges [0, fluid_blocks), then boundary ranges
    LoopState* states;            // [S]
    uint32_t scene_count;
    uint32_t fluid_blocks, boundary_blocks;
    uint32_t n_total, nb_total;   // list strides
    uint32_t pad;
    const uint2* fgrid;           // dense range grids of the scenes' cell boxes (fluid, boundary), or null
    const uint2* bgrid;
};

// Code of a cell for the storage and work orders.  `bits` packs three fields: the low
// byte ib, the bits per axis of a cell inside the scene's cell box; the next byte ob,
// the bits per axis of a cell outside it; the third byte sh, a right shift of the
// coordinates of a cell outside.  Inside the box (its first 2^ib cells along each
// axis) the code is the Morton interleave of the box-relative coordinates; a cell
// outside (particles thrown clear of the scene) gets the bit above an ob-bit Morton
// code of its box-relative coordinates divided by 2^sh (floor) and offset by 2^(ob-1),
// so it follows every cell inside, and a clump of such particles stays together (in
// blocks of 2^sh cells, within 2^(ob-1+sh) cells of the box) instead of wrapping onto
// others every 2^ib cells.  Without a box the coordinates wrap modulo 2^ib.  The scene
// index sits above: max(ib, ob) * 3 + 1 + scene bits.
template <typename R>
__device__ __forceinline__ uint32_t order_code(const SphParams<R>& P, uint32_t scene, int3 q, uint32_t bits) {
    const uint32_t ib = bits & 0xFFu, ob = (bits >> 8) & 0xFFu, sh = bits >> 16;
    const uint32_t w = ib > ob ? ib : ob;
    uint32_t x = (uint32_t)q.x, y = (uint32_t)q.y, z = (uint32_t)q.z;
    if (P.box_dim[0]) {
        x -= P.box_lo[0];
        y -= P.box_lo[1];
        z -= P.box_lo[2];
        if ((x | y | z) >> ib) {
            const uint32_t half = ob ? 1u << (ob - 1) : 0u;
            const uint32_t cx = (uint32_t)((int32_t)x >> sh) + half, cy = (uint32_t)((int32_t)y >> sh) + half,
                           cz = (uint32_t)((int32_t)z >> sh) + half;
            return (scene << (3 * w + 1)) | (1u << (3 * w)) | spread_bits3(cx, ob) | (spread_bits3(cy, ob) << 1) |
                   (spread_bits3(cz, ob) << 2);
        }
    }
    return (scene << (3 * w + 1)) | spread_bits3(x, ib) | (spread_bits3(y, ib) << 1) | (spread_bits3(z, ib) << 2);
}

// A scene's fluid (or boundary) bucket ranges: the batch's bucket table and grid.
template <typename R>
__device__ __forceinline__ CellTable cell_table(const SphParams<R>& P, const uint2* cells, const uint2* grid) {
    CellTable T;
    T.cells = cells + P.bucket_base;
    T.grid = grid ? grid + P.box_base : nullptr;
    T.g = P.grid;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        T.lo[a] = P.box_lo[a];
        T.dim[a] = grid ? P.box_dim[a] : 0u;
    }
    T.prune = false;
    return T;
}

template <typename R>
struct Bits;
template <>
struct Bits<float> {
    __device__ __host__ static unsigned long long of(float v) {
        union {
            float f;
            uint32_t u;
        } x{v};
        return x.u;
    }
    __device__ __host__ static float to(unsigned long long b) {
        union {
            uint32_t u;
            float f;
        } x{(uint32_t)b};
        return x.f;
    }
};
template <>
struct Bits<double> {
    __device__ __host__ static unsigned long long of(double v) {
        union {
            double f;
            unsigned long long u;
        } x{v};
        return x.u;
    }
    __device__ __host__ static double to(unsigned long long b) {
        union {
            unsigned long long u;
            double f;
        } x{b};
        return x.f;
    }
};

// Termination of iteration k (PCISPH/src/sim/Fluid.cpp:283-306 with the minimum and
// maximum iteration counts as options): tested from the min_iterations-th iteration
// on, stop when max dv / rho0 < threshold, and always after max_iterations.
template <typename R>
__device__ __host__ inline bool iteration_is_last(const unsigned long long* max_dv, uint32_t k,
                                                  uint32_t min_it, uint32_t max_it, R rho0, R thr) {
    if (k + 1 >= max_it) return true;
    if (k + 1 < min_it) return false;
#ifdef __HIP_DEVICE_COMPILE__
    return div_rn(Bits<R>::to(max_dv[k]), rho0) < thr;
#else
    return Bits<R>::to(max_dv[k]) / rho0 < thr;
#endif
}

// Iteration k runs unless an earlier one ended the loop.  Iterations after the last
// one find max_dv == 0 (cleared per step) in their predecessor, which reads as
// converged for any positive threshold, so one look back is enough.
template <typename R>
__device__ inline bool iteration_skipped(const LoopState* st, uint32_t k, const SphParams<R>& P) {
    if (k == 0) return false;
    return iteration_is_last<R>(st->max_dv, k - 1, P.min_iterations, P.max_iterations, P.rho0,
                                P.threshold);
}

// The pressure pass of iteration k leaves a scene out when an earlier iteration ended
// its loop, and at iteration 0 when the step starts it from the network: the dual
// pass (k_pressure_dual) runs that iteration for it.
template <typename R>
__device__ inline bool pressure_pass_skipped(const LoopState* st, uint32_t k, const SphParams<R>& P) {
    return (k == 0 && P.warm) || iteration_skipped<R>(st, k, P);
}

// ------------------------------------------------------------- arithmetic modes
template <typename R, bool FAST>
struct Ops;
template <>
struct Ops<float, false> {
    // `/` and __builtin_sqrtf are correctly rounded in HIP's default mode (the
    // v_sqrt_f32 result is corrected by two FMAs); __fsqrt_rn is NOT: without
    // OCML_BASIC_ROUNDED_OPERATIONS it maps to the native (1 ulp) square root.
    __device__ static float div(float a, float b) { return a / b; }
    __device__ static float sqrt(float a) { return __builtin_sqrtf(a); }
};
template <>
struct Ops<float, true> {
    __device__ static float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static float sqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
};
template <>
struct Ops<double, false> {
    __device__ static double div(double a, double b) { return a / b; }
    __device__ static double sqrt(double a) { return ::sqrt(a); }
};

template <typename R>
__device__ __forceinline__ R dist2(R ax, R ay, R az, R bx, R by, R bz, R& dx, R& dy, R& dz) {
    dx = ax - bx;
    dy = ay - by;
    dz = az - bz;
    return dx * dx + dy * dy + dz * dz;
}

// poly6 without normalisation (pcisph.cl:8-16)
template <typename R>
__device__ __forceinline__ R poly6(R r2, R h2) {
    if (r2 > h2) return R(0);
    const R dr = h2 - r2;
    return dr * dr * dr;
}

// A pressure pass's poly6 term of the pair (a, b) added to `sum` (the sums of
// pcisph.cl:238-252 over poly6, pcisph.cl:8-16).  IEEE FP32 and FP64: the reference's
// operations -- r^2, the branch at h^2, (h^2 - r^2)^3 -- one rounding each.  Fast FP32:
// h^2 - r^2 by three fused multiply-adds from h^2, clamped at zero and cubed into the
// sum by one more; the term is the same to within fused rounding and exactly zero
// beyond h, as the branch makes it.  The clamp is taken to [0, 1], which the compiler
// folds into the last multiply-add's output modifier (eight vector operations per
// candidate instead of twelve): it is max(0, h^2 - r^2) because h^2 - r^2 <= h^2 <= 1,
// and the solver runs the fast kernels only where every scene has h <= 1 m (FP32 IEEE
// otherwise).  A NaN distance adds zero there; the pass's max(0, rho* - rho0) drops a
// NaN sum in every mode.
template <typename R, bool FAST>
__device__ __forceinline__ R poly6_add(R sum, const V4<R>& a, const V4<R>& b, R h2) {
    if constexpr (FAST && sizeof(R) == 4) {
        const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
        float e = __builtin_fmaf(-dx, dx, h2);
        e = __builtin_fmaf(-dy, dy, e);
        e = __builtin_fmaf(-dz, dz, e);
        e = __builtin_fminf(__builtin_fmaxf(e, 0.f), 1.f);
        return __builtin_fmaf(e * e, e, sum);
    } else {
        R dx, dy, dz;
        return sum + poly6<R>(dist2<R>(a.x, a.y, a.z, b.x, b.y, b.z, dx, dy, dz), h2);
    }
}

// Akinci cohesion spline (pcisph.cl:48-61)
template <typename R>
__device__ __forceinline__ R surface_tension_kernel(R r, R h, R st_term) {
    if (r <= R(0.5) * h) {
        const R t1 = h - r;
        const R t2 = (t1 * t1 * t1) * (r * r * r);
        return R(2) * t2 - st_term;
    } else if (r <= h) {
        const R t1 = h - r;
        return (t1 * t1 * t1) * (r * r * r);
    }
    return R(0);
}

template <typename R>
__device__ __forceinline__ int3 cell_of4(const V4<R>& p, R cell) {
    return cell_of<R>(p.x, p.y, p.z, cell);
}

template <typename R, typename F>
__device__ __forceinline__ void for_each_candidate(const uint2* __restrict__ cells, int3 c,
                                                   const HashGrid& g, F&& body) {
    for_each_bucket(c, g, [&](uint32_t key) {
        const uint2 r = cells[key];
        for (uint32_t j = r.x; j < r.y; ++j) body(j);
    });
}

// Per-step neighbour lists in chunks of four 24-bit entries, slot-major: entries
// 4c..4c+3 of particle i are packed in the uint3 at [c * stride + i] (pack_chunk), so
// the lanes of a wavefront read consecutive 12-byte chunks and a lane writes its list
// 12 bytes at a time (lists are kept only while every index fits in 24 bits): for
// every fluid particle its
// fluid and boundary candidates within the list radius, for every boundary particle
// its fluid candidates, each in the reference's bucket order.  A particle with more
// neighbours than the capacity gets kListOverflow and walks its buckets.
constexpr uint32_t kListOverflow = 0xFFFFFFFFu;

//
// The pressure pass measures distances at predicted positions x^ but looks up the
// 27 buckets of x (PCISPH/data/kernels/pcisph.cl:238,246).  A candidate j of those
// buckets with |x^_i - x^_j| <= h has |x_i - x_j| <= h + d_i + d_j (d = |x^ - x|),
// so it is in i's list whenever d_i + d_j <= skin h.  disp[k & 1] holds, per bucket,
// the largest d of its fluid particles for iteration k's predicted positions (float
// bits; written by the pass that predicts them, cleared one pass ahead), and
// particle i walks its list in iteration k when d_i + D_i <= 0.9 skin h, D_i the
// largest entry of its 27 buckets (a boundary particle, d = 0, when D_i <= 0.9 skin h).
struct NeighborLists {
    uint32_t* f = nullptr;        // fluid neighbours of fluid particles    [cap_f][n]
    uint32_t* b = nullptr;        // boundary neighbours of fluid particles [cap_b][n]
    uint32_t* bf = nullptr;       // fluid neighbours of boundary particles [cap_bf][nb]
    uint32_t* fcount = nullptr;   // [n]
    uint32_t* bcount = nullptr;   // [n]
    uint32_t* bfcount = nullptr;  // [nb]
    uint32_t* disp[2] = {nullptr, nullptr};  // per-bucket max displacement, by iteration parity [B]
    uint32_t cap_f = 0, cap_b = 0, cap_bf = 0;
    uint32_t on = 0;       // 0: no lists this step (every pass walks its buckets)
    uint32_t buckets = 0;  // entries of each disp table
    uint32_t pad = 0;
};

// Clears a displacement table: the whole grid shares the work.
__device__ __forceinline__ void clear_table(uint32_t* t, uint32_t count) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < count; q += stride) t[q] = 0u;
}

// Records each particle's displacement in its bucket's entry: a segmented maximum
// over the lanes of the wavefront that share the bucket (sorted order keeps them
// adjacent), then one atomic per segment.  Inactive lanes pass key 0xFFFFFFFF.
__device__ __forceinline__ void record_displacement(uint32_t* table, uint32_t key, float d) {
    const int lane = threadIdx.x & 63;
    uint32_t b = __float_as_uint(d);
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t ob = __shfl_up(b, off), ok = __shfl_up(key, off);
        if (lane >= off && ok == key && ob > b) b = ob;
    }
    const uint32_t next = __shfl_down(key, 1);
    if (key != 0xFFFFFFFFu && (lane == 63 || next != key) && b) atomicMax(&table[key], b);
}

// D_i: the largest recorded displacement in the 27 buckets around cell c (all 27
// independent loads issued together).
__device__ __forceinline__ float bucket_displacement(const uint32_t* __restrict__ table, int3 c,
                                                     const HashGrid& g, uint32_t base) {
    uint32_t v[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const uint32_t x = (uint32_t)c.x + (uint32_t)(k % 3 - 1);
        const uint32_t y = (uint32_t)c.y + (uint32_t)((k / 3) % 3 - 1);
        const uint32_t z = (uint32_t)c.z + (uint32_t)(k / 9 - 1);
        v[k] = table[base + hash_key(x, y, z, g)];
    }
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 27; ++k) m = v[k] > m ? v[k] : m;
    return __uint_as_float(m);
}

// Appends to particle i's list, four entries per store (the chunk's tail past the
// list's length is left undefined).  `cap` is a multiple of 4.
// Four list entries (indices below 2^24) in three words, and back.
static_assert(sizeof(uint3) == 12 && alignof(uint3) <= 4, "list chunks are three packed words");
__device__ __forceinline__ uint3 pack_chunk(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    return make_uint3(a | (b << 24), (b >> 8) | (c << 16), (c >> 16) | (d << 8));
}
__device__ __forceinline__ void unpack_chunk(const uint3 w, uint32_t (&e)[4]) {
    e[0] = w.x & 0xFFFFFFu;
    e[1] = (w.x >> 24) | ((w.y & 0xFFFFu) << 8);
    e[2] = (w.y >> 16) | ((w.z & 0xFFu) << 16);
    e[3] = w.z >> 8;
}

// The chunk being filled is kept as one 96-bit word w (w[0] lowest) that each push
// shifts right by 24 bits, entering the new entry at the top: after four pushes it
// holds pack_chunk of them, oldest lowest (three funnel shifts a push, where placing an
// entry by its position in the chunk costs a compare and a select per entry).
struct ListWriter {
    uint32_t* base;
    uint32_t stride, i, cap, len;
    uint32_t w[3];
    __device__ __forceinline__ void shift_in(uint32_t j) {
        w[0] = __builtin_amdgcn_alignbit(w[1], w[0], 24u);
        w[1] = __builtin_amdgcn_alignbit(w[2], w[1], 24u);
        w[2] = __builtin_amdgcn_alignbit(j, w[2], 24u);
    }
    __device__ __forceinline__ void store_chunk() {
        reinterpret_cast<uint3*>(base)[(size_t)(len >> 2) * stride + i] = make_uint3(w[0], w[1], w[2]);
    }
    __device__ __forceinline__ void push(uint32_t j) {
        shift_in(j);
        if ((len & 3u) == 3u && len < cap) store_chunk();
        ++len;
    }
    // the last, partial chunk (shifted down to the bottom); then the length
    // (kListOverflow past the capacity)
    __device__ __forceinline__ uint32_t finish() {
        const uint32_t r = len & 3u;
        if (r && len < cap) {
            for (uint32_t u = r; u < 4u; ++u) shift_in(0u);
            store_chunk();
        }
        return len <= cap ? len : kListOverflow;
    }
};


template <typename T>
__device__ __forceinline__ T block_reduce_sum(T v, T* lds) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    T s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += lds[w];
    return s;
}

// ------------------------------------------------------------- kernels
//
// The step kernels run workgroups of at most kMaxBlockThreads (the solver's
// block_size option), which lets the compiler give a lane up to 256 VGPRs: the force
// initialisation's batched gathers (three float4 per candidate) spill to scratch under
// the 128 a 1024-thread bound allows.
constexpr int kMaxBlockThreads = 512;
//
// Every step kernel starts from its workgroup's BlockRange: the scene (whose
// SphParams and LoopState it uses) and the range of particle indices it covers.
// Indices are batch-wide; a scene's buckets are its own slice of the cell tables.

// initialize (PCISPH/data/kernels/sort_utils.cl:12-23): key = bucket_base +
// hash(cell(pos)), src = index.  `blocks` is the fluid or the boundary part of the
// block table.
// With `canon` (fluid particles stored in the Morton layout) the particle in slot i
// has canonical index canon[i] -- its index in the reference's arrays -- and its key
// goes there, src = the slot it is read from.
template <typename R>
__global__ void k_hash(const SphParams<R>* __restrict__ scenes, const BlockRange* __restrict__ blocks,
                       const V4<R>* __restrict__ pos, uint32_t* __restrict__ keys, uint32_t* __restrict__ src,
                       const uint32_t* __restrict__ canon) {
    const BlockRange rg = blocks[blockIdx.x];
    const uint32_t i = rg.begin + threadIdx.x;
    if (i >= rg.end) return;
    const SphParams<R>& P = scenes[rg.scene];
    const int3 c = cell_of4<R>(pos[i], P.cell);
    const uint32_t at = canon ? canon[i] : i;
    keys[at] = P.bucket_base + hash_key((uint32_t)c.x, (uint32_t)c.y, (uint32_t)c.z, P.grid);
    src[at] = i;
}

// reorder_and_insert_{boundary,fluid}_offsets (sort_utils.cl:25-78): gather by the
// permutation and record bucket start/end at key transitions.
template <typename R>
__global__ void k_reorder(uint32_t n, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ src,
                          const V4<R>* __restrict__ pin, const V4<R>* __restrict__ vin,
                          V4<R>* __restrict__ pout, V4<R>* __restrict__ vout, uint2* __restrict__ cells) {
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    const uint32_t s = src[d];
    pout[d] = pin[s];
    if (vin) vout[d] = vin[s];
    const uint32_t k = keys[d];
    if (d == 0 || k != keys[d - 1]) cells[k].x = d;
    if (d == n - 1 || k != keys[d + 1]) cells[k].y = d + 1;
}

// The dense range grid of scene blockIdx.y's cell box (CellTable): each box cell's
// bucket range, copied from the bucket table after a sort.  With `pos` (the positions
// the table's ranges index), an entry whose bucket holds only particles at least
// kFarCells cells away along some axis gets kFarBucket; cells with a coordinate of
// magnitude 2^15 or more, buckets of more than kFarScan particles and buckets with a NaN
// position (whose terms are NaN at any distance) are never flagged.
constexpr uint32_t kFarScan = 64;

template <typename R>
__global__ void k_fill_box(const SphParams<R>* __restrict__ scenes, const uint2* __restrict__ cells,
                           uint2* __restrict__ grid, const V4<R>* __restrict__ pos) {
    const SphParams<R>& P = scenes[blockIdx.y];
    const uint32_t dx = P.box_dim[0], dy = P.box_dim[1], total = dx * dy * P.box_dim[2];
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
        const uint32_t rx = q % dx, rest = q / dx, ry = rest % dy, rz = rest / dy;
        const uint32_t cx = P.box_lo[0] + rx, cy = P.box_lo[1] + ry, cz = P.box_lo[2] + rz;
        uint2 r = cells[P.bucket_base + hash_key(cx, cy, cz, P.grid)];
        const int32_t n[3] = {(int32_t)cx, (int32_t)cy, (int32_t)cz};
        bool far = pos && r.y > r.x && r.y - r.x <= kFarScan;
#pragma unroll
        for (int a = 0; a < 3; ++a) far = far && n[a] > -32768 && n[a] < 32768;
        for (uint32_t j = r.x; far && j < r.y; ++j) {
            const V4<R> o = pos[j];
            const int3 m = cell_of4<R>(o, P.cell);
            const int64_t d[3] = {(int64_t)m.x - n[0], (int64_t)m.y - n[1], (int64_t)m.z - n[2]};
            bool away = false;
#pragma unroll
            for (int a = 0; a < 3; ++a) away = away || d[a] >= kFarCells || d[a] <= -kFarCells;
            far = away && o.x == o.x && o.y == o.y && o.z == o.z;
        }
        if (far) r.x |= kFarBucket;
        grid[P.box_base + q] = r;
    }
}

// kQuietCell on every cell of scene blockIdx.y's box whose 27 neighbours lie in the
// box and whose entries are all empty or kFarBucket (after k_fill_box with positions).
// Only the end words' top bits change, and every range reader masks them off, so cells
// read while their neighbours are being marked still see the right ranges.
template <typename R>
__global__ void k_mark_quiet(const SphParams<R>* __restrict__ scenes, uint2* __restrict__ grid) {
    const SphParams<R>& P = scenes[blockIdx.y];
    const uint32_t dx = P.box_dim[0], dy = P.box_dim[1], dz = P.box_dim[2], total = dx * dy * dz;
    uint2* G = grid + P.box_base;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
        const uint32_t rx = q % dx, rest = q / dx, ry = rest % dy, rz = rest / dy;
        if (rx < 1 || ry < 1 || rz < 1 || rx + 1 >= dx || ry + 1 >= dy || rz + 1 >= dz) continue;
        bool quiet = true;
        for (int o = 0; o < 27 && quiet; ++o) {
            const uint2 e = G[(rx + o % 3 - 1) + dx * ((ry + (o / 3) % 3 - 1) + dy * (rz + o / 9 - 1))];
            quiet = (e.x & kFarBucket) || (e.x & ~kFarBucket) == (e.y & ~kQuietCell);
        }
        if (quiet) G[q].y |= kQuietCell;
    }
}

// Cell extents of every scene's fluid and boundary particles: ext[12 s ..] holds the
// fluid minimum (x, y, z) and maximum, then the boundary minimum and maximum, as
// int32 folded with atomics (the caller sets minima to INT32_MAX, maxima to INT32_MIN).
template <typename R>
__global__ __launch_bounds__(kMaxBlockThreads) void k_extents(const SphParams<R>* __restrict__ scenes,
                                                            const BlockRange* __restrict__ blocks,
                                                            const V4<R>* __restrict__ pos,
                                                            const V4<R>* __restrict__ bpos, int* __restrict__ ext) {
    const BlockRange rg = blocks[blockIdx.x];
    const uint32_t i = rg.begin + threadIdx.x;
    const bool boundary = (rg.flags & kRangeBoundary) != 0;
    int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    if (i < rg.end) {
        const int3 c = cell_of4<R>(boundary ? bpos[i] : pos[i], scenes[rg.scene].cell);
        lo[0] = hi[0] = c.x;
        lo[1] = hi[1] = c.y;
        lo[2] = hi[2] = c.z;
    }
    // wavefront, then workgroup reduction: one atomic per value and workgroup
    __shared__ int s_ext[6][kMaxBlockThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            lo[a] = min(lo[a], __shfl_xor(lo[a], off));
            hi[a] = max(hi[a], __shfl_xor(hi[a], off));
        }
        if (lane == 0) {
            s_ext[a][wave] = lo[a];
            s_ext[3 + a][wave] = hi[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int q = threadIdx.x;
        int v = q < 3 ? INT32_MAX : INT32_MIN;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v = q < 3 ? min(v, s_ext[q][w]) : max(v, s_ext[q][w]);
        int* e = ext + 12 * rg.scene + (boundary ? 6 : 0) + q;
        if (q < 3) atomicMin(e, v);
        else atomicMax(e, v);
    }
}

// Wave-wide maximum of a non-negative float, folded into *dst (uint bits order =
// float order for x >= 0).  Every lane of the wave must call it.
__device__ __forceinline__ void wave_max_to(unsigned int* dst, float v) {
    unsigned int b = __float_as_uint(v);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned int o = __shfl_xor(b, off);
        b = o > b ? o : b;
    }
    if ((threadIdx.x & 63) == 0 && b) atomicMax(dst, b);
}

// Workgroup-wide maxima of up to four 64-bit values, folded into dst[q] by thread 0.
template <int Q>
__device__ __forceinline__ void block_max_to(unsigned long long (&m)[Q], unsigned long long* const (&dst)[Q],
                                             unsigned long long (*lds)[8]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(m[q], off);
            m[q] = o > m[q] ? o : m[q];
        }
        if (lane == 0) lds[q][wave] = m[q];
    }
    __syncthreads();
    if (threadIdx.x < Q) {
        unsigned long long v = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v = lds[threadIdx.x][w] > v ? lds[threadIdx.x][w] : v;
        // the atomic is skipped when the stored maximum is already at least v: a
        // device-coherent read returns the step's reset value (ordered before this
        // launch by the stream) or a later, larger one, so a skipped v never exceeds it
        if (v && v > *(volatile unsigned long long*)dst[threadIdx.x]) atomicMax(dst[threadIdx.x], v);
    }
}

template <typename R>
__device__ __forceinline__ float displacement(const V4<R>& a, const V4<R>& b) {
    const float dx = (float)(a.x - b.x), dy = (float)(a.y - b.y), dz = (float)(a.z - b.z);
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

// Batched gathers.  A neighbour walk is split into load(j) -> data and
// use(data, j): a batch of candidates is loaded before any is used, so each lane
// keeps that many independent loads in flight instead of one; use() still sees the
// candidates in order, so sums are unchanged.  Measured best: 4 for bucket ranges
// (every pass that walks buckets, the list-building density pass included), 8 for
// neighbour lists (index chunks first, then data).
constexpr int kGatherBatch = 4;
constexpr int kListBatch = 8;
constexpr int kForceBatch = 8;  // force init: three float4 per candidate

template <typename Load, typename Use>
__device__ __forceinline__ void gather_range(uint32_t start, uint32_t end, Load&& load, Use&& use) {
    using T = decltype(load(0u));
    uint32_t j = start;
    for (; j + kGatherBatch <= end; j += kGatherBatch) {
        T d[kGatherBatch];
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u) d[u] = load(j + u);
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u) use(d[u], j + u);
    }
    for (; j < end; ++j) use(load(j), j);
}

// The three ranges of a bucket row gathered as one sequence (r0, then r1, then r2):
// batches of U loads run across bucket boundaries, so a row of ~24 candidates is a
// few full batches instead of three short ranges each ending in a remainder loop
// (measured: the pressure pass 466 -> 362 us at 1M with U = 4).
template <int U, typename Load, typename Use>
__device__ __forceinline__ void gather_row(uint2 r0, uint2 r1, uint2 r2, Load&& load, Use&& use) {
    using T = decltype(load(0u));
    const uint32_t n0 = r0.y - r0.x, n01 = n0 + (r1.y - r1.x), n = n01 + (r2.y - r2.x);
    auto at = [&](uint32_t k) { return k < n0 ? r0.x + k : (k < n01 ? r1.x + (k - n0) : r2.x + (k - n01)); };
    uint32_t k = 0;
    for (; k + U <= n; k += U) {
        uint32_t jj[U];
        T d[U];
#pragma unroll
        for (int u = 0; u < U; ++u) jj[u] = at(k + u);
#pragma unroll
        for (int u = 0; u < U; ++u) d[u] = load(jj[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) use(d[u], jj[u]);
    }
    for (; k < n; ++k) {
        const uint32_t j = at(k);
        use(load(j), j);
    }
}

// The 27 buckets around cell c: a row's three bucket ranges are loaded before any of
// them is walked, then gathered as one sequence.
// `full`, when given, accumulates the candidate count before far buckets are pruned.
template <typename R, int U = kGatherBatch, typename Load, typename Use>
__device__ __forceinline__ void walk_buckets(const CellTable& T, int3 c, Load&& load, Use&& use,
                                             uint32_t* full = nullptr) {
    for_each_range_row(
        T, c, [&](uint2 r0, uint2 r1, uint2 r2) { gather_row<U>(r0, r1, r2, load, use); }, full);
}

// Whether a walk may read far buckets as empty (CellTable): its evaluation points lie
// within d of the walking particle's position and its terms vanish beyond rho.  False
// for a NaN d.
template <typename R>
__device__ __forceinline__ bool far_prunable(float d, float rho, R cell) {
    return d + rho <= kFarReach * (float)cell;
}

// One list chunk; NT: a non-temporal load, so the chunks streaming through a pass
// do not push out of L2 the lines its gathers and the next pass reuse.
template <bool NT>
__device__ __forceinline__ uint3 list_chunk(const uint3* p) {
    if constexpr (NT) {
        typedef unsigned int u32x3 __attribute__((ext_vector_type(3)));
        const u32x3 v = __builtin_nontemporal_load(reinterpret_cast<const u32x3*>(p));
        return make_uint3(v.x, v.y, v.z);
    } else {
        return *p;
    }
}

// Particle i's list of `len` entries (chunks of four, slot-major), U loads in flight
// per lane (U a multiple of 4); NT: chunks read with non-temporal loads.
template <int U = kListBatch, bool NT = false, typename Load, typename Use>
__device__ __forceinline__ void walk_list(const uint32_t* __restrict__ list, uint32_t len, uint32_t stride,
                                          uint32_t i, Load&& load, Use&& use) {
    static_assert(U % 4 == 0, "list batches are whole chunks");
    using T = decltype(load(0u));
    const uint3* q = reinterpret_cast<const uint3*>(list) + i;
    uint32_t k = 0;
    for (; k + U <= len; k += U) {
        uint32_t jj[U];
        T d[U];
#pragma unroll
        for (int c = 0; c < U / 4; ++c) {
            uint32_t e[4];
            unpack_chunk(list_chunk<NT>(q + (size_t)(k / 4 + c) * stride), e);
            jj[4 * c] = e[0];
            jj[4 * c + 1] = e[1];
            jj[4 * c + 2] = e[2];
            jj[4 * c + 3] = e[3];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) d[u] = load(jj[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) use(d[u], jj[u]);
    }
    for (; k < len; k += 4) {
        uint32_t jj[4];
        unpack_chunk(list_chunk<NT>(q + (size_t)(k / 4) * stride), jj);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k + u < len) use(load(jj[u]), jj[u]);
    }
}

// Particle i's list if it has one, else its 27 buckets (same order, same non-zero terms).
template <typename R, int U = kListBatch, bool NT = false, typename Load, typename Use>
__device__ __forceinline__ void walk(const uint32_t* __restrict__ list, uint32_t len, uint32_t stride, uint32_t i,
                                     const CellTable& T, int3 c, Load&& load, Use&& use) {
    if (len != kListOverflow) walk_list<U, NT>(list, len, stride, i, load, use);
    else walk_buckets<R>(T, c, load, use);
}

// A boundary particle's share of the density pass: its pressure reset
// (boundary_pressure_initialization, pcisph.cl:114-119) and, where pressure passes test
// their lists, its list of fluid neighbours.
template <typename R>
__device__ __forceinline__ void density_boundary_particle(const Batch<R>& B, const SphParams<R>& P, uint32_t b,
                                                          const CellTable& FT, const V4<R>* __restrict__ bpos,
                                                          const V4<R>* __restrict__ pos, R* __restrict__ bp,
                                                          V4<R>* __restrict__ bxf, const NeighborLists& NL) {
    bp[b] = 0;
    V4<R> o = bpos[b];
    o.w = 0;
    bxf[b] = o;
    if (NL.on && P.disp_limit > 0.f) {  // read only by pressure passes that test their lists
        const R lr2 = (R)P.list_r2;
        const int3 c = cell_of4<R>(o, P.cell);
        ListWriter w{NL.bf, B.nb_total, b, NL.cap_bf, 0, {}};
        walk_buckets<R>(FT, c, [&](uint32_t j) { return pos[j]; },
                        [&](const V4<R>& q, uint32_t j) {
                            R dx, dy, dz;
                            if (dist2<R>(o.x, o.y, o.z, q.x, q.y, q.z, dx, dy, dz) <= lr2) w.push(j);
                        });
        NL.bfcount[b] = w.finish();
    }
}

// FP32 density passes are built for 8 wavefronts per SIMD (64 VGPRs: a 24-byte spill,
// measured 5-10% faster than 7 wavefronts at 68 VGPRs); FP64 keeps its registers.
template <typename R>
constexpr int kDensityWavesPerEu = sizeof(R) == 4 ? 8 : 1;

// update_density (pcisph.cl:65-92), plus per-particle candidate / neighbour counts,
// plus this step's neighbour lists (every candidate within the list radius, in the
// reference's bucket order).  Boundary ranges reset the boundary pressures
// (boundary_pressure_initialization, pcisph.cl:114-119) and record their fluid
// neighbours.
template <typename R, bool FAST>
__global__ __launch_bounds__(kMaxBlockThreads) __attribute__((amdgpu_waves_per_eu(kDensityWavesPerEu<R>)))
void k_density(Batch<R> B, const uint2* __restrict__ bcells, const V4<R>* __restrict__ bpos,
                          const uint2* __restrict__ fcells, const V4<R>* __restrict__ pos,
                          V4<R>* __restrict__ xr, uint4* __restrict__ counts, R* __restrict__ bp,
                          V4<R>* __restrict__ bxf, NeighborLists NL) {
    __shared__ unsigned long long lds[16];
    const BlockRange rg = B.blocks[blockIdx.x];
    const SphParams<R> P = B.scenes[rg.scene];
    LoopState* st = B.states + rg.scene;
    const CellTable FT = cell_table(P, fcells, B.fgrid);
    const uint32_t t = rg.begin + threadIdx.x;
    const bool in = t < rg.end;
    const bool lists = NL.on != 0;
    const R lr2 = (R)P.list_r2;
    if (lists) clear_table(NL.disp[0], NL.buckets);  // filled by the force pass
    unsigned long long cf_sum = 0, cb_sum = 0;
    if (!(rg.flags & kRangeBoundary)) {
        if (in) {
            const uint32_t i = t;
            const V4<R> self = pos[i];
            const int3 c = cell_of4<R>(self, P.cell);
            R rho = 0;
            uint32_t cf = 0, cb = 0, nf = 0, nb = 0;
            ListWriter wb{NL.b, B.n_total, t, NL.cap_b, 0, {}}, wf{NL.f, B.n_total, t, NL.cap_f, 0, {}};
            if (P.nb) {
                // terms vanish beyond h; list entries beyond the list radius
                CellTable BT = cell_table(P, bcells, B.bgrid);
                BT.prune = far_prunable<R>(0.f, sqrtf(fmaxf((float)P.h2, lists ? (float)lr2 : 0.f)), P.cell);
                walk_buckets<R>(
                    BT, c, [&](uint32_t j) { return bpos[j]; },
                    [&](const V4<R>& o, uint32_t j) {
                        R dx, dy, dz;
                        const R r2 = dist2<R>(self.x, self.y, self.z, o.x, o.y, o.z, dx, dy, dz);
                        rho += poly6<R>(r2, P.h2);
                        nb += !(r2 > P.h2);
                        if (lists && r2 <= lr2) wb.push(j);
                    },
                    &cb);
            }
            walk_buckets<R>(FT, c, [&](uint32_t j) { return pos[j]; },
                            [&](const V4<R>& o, uint32_t j) {
                                R dx, dy, dz;
                                const R r2 = dist2<R>(self.x, self.y, self.z, o.x, o.y, o.z, dx, dy, dz);
                                rho += poly6<R>(r2, P.h2);
                                ++cf;
                                nf += !(r2 > P.h2);
                                if (lists && r2 <= lr2) wf.push(j);
                            });
            rho *= P.density_scale;
            V4<R> out = self;
            out.w = rho;
            xr[i] = out;
            if (counts) counts[i] = make_uint4(cf, cb, nf, nb);
            if (lists) {
                NL.bcount[t] = P.nb ? wb.finish() : 0u;
                NL.fcount[t] = wf.finish();
            }
            cf_sum = cf;
            cb_sum = cb;
        }
    } else if (in) {
        density_boundary_particle<R>(B, P, t, FT, bpos, pos, bp, bxf, NL);
    }
    const unsigned long long a = block_reduce_sum(cf_sum, lds);
    __syncthreads();
    const unsigned long long b = block_reduce_sum(cb_sum, lds + 8);
    if (threadIdx.x == 0 && (a || b)) {
        atomicAdd(&st->fluid_candidates, a);
        atomicAdd(&st->boundary_candidates, b);
    }
}

// update_normal (pcisph.cl:94-112); fluid ranges only
template <typename R, bool FAST>
__global__ __launch_bounds__(kMaxBlockThreads) void k_normal(Batch<R> B, const uint2* __restrict__ fcells, const V4<R>* __restrict__ xr,
                         V4<R>* __restrict__ nrm, NeighborLists NL) {
    using O = Ops<R, FAST>;
    const BlockRange rg = B.blocks[blockIdx.x];
    const uint32_t t = rg.begin + threadIdx.x;
    if (t >= rg.end) return;
    const uint32_t i = t;
    const SphParams<R> P = B.scenes[rg.scene];
    const V4<R> self = xr[i];
    const int3 c = cell_of4<R>(self, P.cell);
    R nx = 0, ny = 0, nz = 0;
    auto body = [&](const V4<R>& o, uint32_t) {
        R dx, dy, dz;
        const R r2 = dist2<R>(self.x, self.y, self.z, o.x, o.y, o.z, dx, dy, dz);
        if (r2 > P.h2) return;  // poly6_d1 is exactly zero there (pcisph.cl:20-21)
        const R w = (r2 - P.h2) * (r2 - P.h2);
        nx += O::div(w * dx, o.w);
        ny += O::div(w * dy, o.w);
        nz += O::div(w * dz, o.w);
    };
    const uint32_t len = NL.on ? NL.fcount[t] : kListOverflow;
    walk<R>(NL.f, len, B.n_total, t, cell_table(P, fcells, B.fgrid), c, [&](uint32_t j) { return xr[j]; }, body);
    V4<R> out;
    out.x = nx * P.normal_scale;
    out.y = ny * P.normal_scale;
    out.z = nz * P.normal_scale;
    out.w = 0;
    nrm[i] = out;
}

// symplectic Euler of update_position_and_velocity (pcisph.cl:175-193)
template <typename R, bool FAST>
__device__ __forceinline__ void integrate(const SphParams<R>& P, const V4<R>& x, const V4<R>& v, R fx, R fy,
                                          R fz, V4<R>& xo, V4<R>& vo) {
    using O = Ops<R, FAST>;
    const R ax = O::div(fx, P.m), ay = O::div(fy, P.m), az = O::div(fz, P.m);
    vo.x = v.x + ax * P.dt;
    vo.y = v.y + ay * P.dt;
    vo.z = v.z + az * P.dt;
    vo.w = 0;
    xo.x = x.x + vo.x * P.dt;
    xo.y = x.y + vo.y * P.dt;
    xo.z = x.z + vo.z * P.dt;
    xo.w = 0;
}

// force_initialization (pcisph.cl:121-173) fused with the first prediction of the
// correction loop (pcisph.cl:175-193 with F_p = 0, Fluid.cpp:253-260); fluid ranges.
// With features requested it also gathers the warm-start network's inputs; the
// velocity's w component carries the previous step's converged pressure.  The
// pressure-force reset to zero is not stored: that array is read only after the
// step, and the loop's last pressure-force pass always writes it.
template <typename R, bool FAST>
__global__ __launch_bounds__(kMaxBlockThreads) void k_force_init(Batch<R> B, const uint2* __restrict__ fcells, const V4<R>* __restrict__ xr,
                             const V4<R>* __restrict__ vel, const V4<R>* __restrict__ nrm,
                             V4<R>* __restrict__ fo, R* __restrict__ p, V4<R>* __restrict__ xp,
                             FeatureArgs F, NeighborLists NL) {
    using O = Ops<R, FAST>;
    const BlockRange rg = B.blocks[blockIdx.x];
    const SphParams<R> P = B.scenes[rg.scene];
    const uint32_t t = rg.begin + threadIdx.x;
    float moved = 0.f;
    uint32_t bucket = 0xFFFFFFFFu;
    if (t < rg.end) {
        const uint32_t i = t;
        const V4<R> self = xr[i];
        const V4<R> sv = vel[i];
        const V4<R> sn = nrm[i];
        const int3 c = cell_of4<R>(self, P.cell);
        bucket = P.bucket_base + hash_key((uint32_t)c.x, (uint32_t)c.y, (uint32_t)c.z, P.grid);
        R vx = 0, vy = 0, vz = 0, cx = 0, cy = 0, cz = 0, ux = 0, uy = 0, uz = 0;
        float a_rho = 0.f, a_p = 0.f, a_div = 0.f, a_n = 0.f;
        const bool want = F.feat != nullptr && (F.all || P.warm);
        struct Cand {
            V4<R> o, ov, on;
        };
        auto load = [&](uint32_t j) { return Cand{xr[j], vel[j], nrm[j]}; };
        auto body = [&](const Cand& cd, uint32_t) {
            const V4<R>& o = cd.o;
            R dx, dy, dz;
            const R r2 = dist2<R>(self.x, self.y, self.z, o.x, o.y, o.z, dx, dy, dz);
            if (r2 > P.h2) return;  // every term below vanishes for dist >= h
            const R dist = O::sqrt(r2);
            const V4<R>& ov = cd.ov;
            if (want) {
                a_n += 1.f;
                a_rho += (float)o.w;
                a_p += (float)ov.w;
                if (dist > R(0.0001f)) {
                    const float dr = (float)(P.h - dist);
                    a_div += (float)(((ov.x - sv.x) * dx + (ov.y - sv.y) * dy + (ov.z - sv.z) * dz) / dist) *
                             dr * dr / (float)o.w;
                }
            }
            if (o.w > R(0.0001f)) {
                const R k = O::div(dist > P.h ? R(0) : P.h - dist, o.w);
                vx += (ov.x - sv.x) * k;
                vy += (ov.y - sv.y) * k;
                vz += (ov.z - sv.z) * k;
            }
            if (dist > R(0.0001f) && dist < P.h) {
                const V4<R>& on = cd.on;
                const R corr = O::div(R(2) * P.rho0, self.w + o.w);
                const R stk = surface_tension_kernel<R>(dist, P.h, P.st_term);
                const R cs = corr * stk;
                cx += cs * O::div(dx, dist);
                cy += cs * O::div(dy, dist);
                cz += cs * O::div(dz, dist);
                ux += corr * (sn.x - on.x);
                uy += corr * (sn.y - on.y);
                uz += corr * (sn.z - on.z);
            }
        };
        const uint32_t len = NL.on ? NL.fcount[t] : kListOverflow;
        walk<R, kForceBatch>(NL.f, len, B.n_total, t, cell_table(P, fcells, B.fgrid), c, load, body);
        const R ox = ((R(0) + vx * P.visc_scale) + (cx * P.coh_scale + ux * P.curv_scale));
        const R oy = ((P.gravity_force + vy * P.visc_scale) + (cy * P.coh_scale + uy * P.curv_scale));
        const R oz = ((R(0) + vz * P.visc_scale) + (cz * P.coh_scale + uz * P.curv_scale));
        V4<R> f;
        f.x = ox;
        f.y = oy;
        f.z = oz;
        f.w = 0;
        fo[i] = f;
        p[i] = 0;
        V4<R> xo, vo;
        integrate<R, FAST>(P, self, sv, ox + R(0), oy + R(0), oz + R(0), xo, vo);
        xp[i] = xo;
        moved = displacement<R>(xo, self);
        if (want) {
            const float rho0 = (float)P.rho0;
            const uint4 cnt = F.counts[i];
            const float inv = a_n > 0.f ? 1.f / a_n : 0.f;
            const float speed = sqrtf((float)(sv.x * sv.x + sv.y * sv.y + sv.z * sv.z));
            float4* out = reinterpret_cast<float4*>(F.feat + (size_t)i * 8);
            out[0] = make_float4(((float)self.w / rho0 - 1.f) * 100.f, (float)cnt.z / 32.f, (float)sv.w / P.p_ref,
                                 speed * (float)P.dt / (float)P.h * 100.f);
            out[1] = make_float4((a_rho * inv / rho0 - 1.f) * 100.f, a_p * inv / P.p_ref, (float)cnt.w / 32.f,
                                 a_div * P.div_scale);
        }
    }
    if (NL.on && P.disp_limit > 0.f) record_displacement(NL.disp[0], bucket, moved);
}

// update_pressure (pcisph.cl:212-263): fluid mode on fluid ranges, boundary mode on
// boundary ranges (the two launches of Fluid.cpp:263-280 are independent).
// Also publishes x + p/rho^2 for the force pass and the per-iteration max dv.
// Distances are taken at the predicted positions, so a wavefront walks the step's
// lists (built at x with radius (1+skin)h) only where every lane's displacement test
// passes (NeighborLists), else its 27 buckets.
// One workgroup's range `range` of the block table (k_pressure, k_correction_loop).
template <typename R, bool FAST>
__device__ __forceinline__ void pressure_range(const Batch<R>& B, uint32_t range, uint32_t iter,
                                               const uint2* __restrict__ bcells, const V4<R>* __restrict__ bpos,
                                               const R* __restrict__ binit, const uint2* __restrict__ fcells,
                                               const V4<R>* __restrict__ xr, const V4<R>* __restrict__ xp,
                                               R* __restrict__ dv, R* __restrict__ p, V4<R>* __restrict__ xf,
                                               R* __restrict__ bp, V4<R>* __restrict__ bxf, const NeighborLists& NL) {
    using O = Ops<R, FAST>;
    __shared__ unsigned long long lds[2][8];
    __shared__ unsigned int s_list_waves;
    const BlockRange rg = B.blocks[range];
    const SphParams<R> P = B.scenes[rg.scene];
    LoopState* st = B.states + rg.scene;
    if (pressure_pass_skipped<R>(st, iter, P)) return;
    if (threadIdx.x == 0) s_list_waves = 0;
    const CellTable FT = cell_table(P, fcells, B.fgrid);
    const uint32_t t = rg.begin + threadIdx.x;
    const bool boundary = (rg.flags & kRangeBoundary) != 0;
    const uint32_t i = t;
    // lists for this wavefront if every lane's pairs are all in its list; a scene
    // fewer than 1% of whose fluid wavefronts could use them in the step's passes so
    // far stops testing from the third pass on (displacements hardly change within a
    // step; boundary wavefronts far from the fluid pass trivially and are not counted)
    const uint32_t waves = (P.n + 63) / 64;
    const bool try_lists = NL.on && P.disp_limit > 0.f &&
                           (iter < 2 || 100ull * st->list_waves >= (unsigned long long)waves * iter);
    bool ok = true;
    V4<R> self{}, pr{};
    int3 c{};
    if (t < rg.end) {
        self = boundary ? bpos[i] : xr[i];
        pr = boundary ? self : xp[i];
        c = cell_of4<R>(self, P.cell);
        if (try_lists) {
            const float d = boundary ? 0.f : displacement<R>(pr, self);
            ok = d + bucket_displacement(NL.disp[iter & 1], c, P.grid, P.bucket_base) <= P.disp_limit;
        }
    }
    const bool lists = try_lists && __all(ok);
    __syncthreads();
    if (lists && !boundary && (threadIdx.x & 63) == 0) atomicAdd(&s_list_waves, 1u);
    unsigned long long m[2] = {0, 0};  // max dv, max expansion
    if (!boundary) {
        if (t < rg.end) {
            R pd = 0;
            if (P.nb) {
                const uint32_t len = lists ? NL.bcount[t] : kListOverflow;
                CellTable BT = cell_table(P, bcells, B.bgrid);
                BT.prune = far_prunable<R>(displacement<R>(pr, self), (float)P.h, P.cell);
                if (len != kListOverflow || !(BT.prune && quiet_cell(BT, c))) {
                    walk<R>(NL.b, len, B.n_total, t, BT, c, [&](uint32_t j) { return bpos[j]; },
                            [&](const V4<R>& o, uint32_t) { pd = poly6_add<R, FAST>(pd, pr, o, P.h2); });
                }
            }
            const uint32_t len = lists ? NL.fcount[t] : kListOverflow;
            walk<R>(NL.f, len, B.n_total, t, FT, c, [&](uint32_t j) { return xp[j]; },
                    [&](const V4<R>& o, uint32_t) { pd = poly6_add<R, FAST>(pd, pr, o, P.h2); });
            pd *= P.density_scale;
            const R v = fmax(R(0), pd - P.rho0);
            // expansion counts only where the step began at or above rest density: free-
            // surface particles sit below rho0 whatever the pressure
            if (self.w >= P.rho0) m[1] = Bits<R>::of(fmax(R(0), P.rho0 - pd));
            dv[i] = v;
            R pi = p[i];
            if (v != R(0)) {
                pi += v * P.delta;
                p[i] = pi;
            }
            V4<R> out = self;
            out.w = O::div(pi, self.w * self.w);
            xf[i] = out;
            m[0] = Bits<R>::of(v);
        }
    } else if (t < rg.end) {
        const uint32_t b = t;
        R pd = 0;
        pd += binit[b];
        const uint32_t len = lists ? NL.bfcount[b] : kListOverflow;
        walk<R>(NL.bf, len, B.nb_total, b, FT, c, [&](uint32_t j) { return xp[j]; },
                [&](const V4<R>& o, uint32_t) { pd = poly6_add<R, FAST>(pd, self, o, P.h2); });
        pd *= P.density_scale;
        const R v = fmax(R(0), pd - P.rho0);
        if (v != R(0)) {
            const R pb = bp[b] + v * P.delta;
            bp[b] = pb;
            V4<R> out = self;
            out.w = O::div(pb, P.rho0_sq);
            bxf[b] = out;
        }
    }
    unsigned long long* const dst[2] = {&st->max_dv[iter], &st->max_exp[iter]};
    block_max_to<2>(m, dst, lds);
    if (threadIdx.x == 0 && s_list_waves) atomicAdd(&st->list_waves, s_list_waves);
}

template <typename R, bool FAST>
__global__ __launch_bounds__(kMaxBlockThreads) void k_pressure(Batch<R> B, uint32_t iter, const uint2* __restrict__ bcells,
                           const V4<R>* __restrict__ bpos, const R* __restrict__ binit,
                           const uint2* __restrict__ fcells, const V4<R>* __restrict__ xr,
                           const V4<R>* __restrict__ xp, R* __restrict__ dv, R* __restrict__ p,
                           V4<R>* __restrict__ xf, R* __restrict__ bp, V4<R>* __restrict__ bxf, NeighborLists NL) {
    if (NL.on) clear_table(NL.disp[(iter + 1) & 1], NL.buckets);  // for the next prediction
    pressure_range<R, FAST>(B, blockIdx.x, iter, bcells, bpos, binit, fcells, xr, xp, dv, p, xf, bp, bxf, NL);
}

// ------------------------------------------------------- LDS-staged pressure pass
//
// The fluid ranges of update_pressure (pcisph.cl:212-263) with the candidates staged
// in LDS.  A workgroup of 256 fluid slots in the Morton layout covers a compact
// block of ~32 cells; the union of its lanes' 27-bucket neighbourhoods is ~150
// buckets.  The workgroup gathers those bucket ids into an LDS hash set, loads each
// bucket's predicted positions (and boundary positions) into LDS once, and every lane
// then walks its 27 buckets in the reference's order from LDS -- the same candidates,
// values and summation order as the global walk, without the per-lane L2 misses.  A
// workgroup whose neighbourhood does not fit walks global memory instead.
// The tile stays below 64 KB of LDS: a workgroup asking for more was measured to run
// alone on the device (each wave busy ~28 us, the launch ~64 x that).
constexpr int kTileThreads = 256;
constexpr uint32_t kTileSlotBits = 9;
constexpr uint32_t kTileSlots = 1u << kTileSlotBits;  // hash-set slots
constexpr uint32_t kTileFluid = 1792;                  // staged fluid candidates
constexpr uint32_t kTileBoundary = 1024;
constexpr uint32_t kTileEmpty = 0xFFFFFFFFu;

struct PressureTile {
    uint32_t key[kTileSlots];
    uint32_t fslot[kTileSlots];  // staged fluid range of the bucket: start | count << 16
    uint32_t bslot[kTileSlots];  // staged boundary range
    uint32_t fsrc[kTileSlots];   // first particle of the bucket in global memory
    uint32_t bsrc[kTileSlots];
    float4 fx[kTileFluid];
    float4 bx[kTileBoundary];
    uint32_t part[2][kTileThreads / 64];
    uint32_t overflow;
};

__device__ __forceinline__ uint32_t tile_hash(uint32_t key) { return (key * 2654435761u) >> (32 - kTileSlotBits); }

__device__ __forceinline__ void tile_insert(PressureTile& T, uint32_t key) {
    uint32_t s = tile_hash(key);
    for (uint32_t probe = 0; probe < kTileSlots; ++probe) {
        const uint32_t seen = T.key[s];
        if (seen == key) return;
        if (seen == kTileEmpty) {
            const uint32_t prev = atomicCAS(&T.key[s], kTileEmpty, key);
            if (prev == kTileEmpty || prev == key) return;
        }
        s = (s + 1) & (kTileSlots - 1);
    }
    T.overflow = 1u;  // the set is full
}

__device__ __forceinline__ uint32_t tile_find(const PressureTile& T, uint32_t key) {
    uint32_t s = tile_hash(key);
    for (uint32_t probe = 0; probe < kTileSlots && T.key[s] != key; ++probe) s = (s + 1) & (kTileSlots - 1);
    return s;
}

// Copies `total` staged elements: element q belongs to the last slot whose staged
// start is <= q (starts are non-decreasing over slots; an empty slot never ends a run
// of equal starts), so every thread copies q = threadIdx.x, +256, ... independently
// and the workgroup keeps many loads in flight.
__device__ __forceinline__ void tile_copy(const uint32_t* __restrict__ slot, const uint32_t* __restrict__ src,
                                          float4* __restrict__ dst, const float4* __restrict__ from,
                                          uint32_t total) {
    for (uint32_t q = threadIdx.x; q < total; q += kTileThreads) {
        uint32_t lo = 0, hi = kTileSlots;  // first slot with start > q
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if ((slot[mid] & 0xFFFFu) <= q) lo = mid + 1;
            else hi = mid;
        }
        const uint32_t s = lo - 1;
        dst[q] = from[src[s] + (q - (slot[s] & 0xFFFFu))];
    }
}

// Exclusive workgroup scan of two counters (fluid, boundary); totals returned.
__device__ __forceinline__ uint2 tile_scan(PressureTile& T, uint32_t a, uint32_t b, uint2& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t xa = a, xb = b;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t ya = __shfl_up(xa, off), yb = __shfl_up(xb, off);
        if (lane >= off) xa += ya, xb += yb;
    }
    if (lane == 63) T.part[0][wave] = xa, T.part[1][wave] = xb;
    __syncthreads();
    uint32_t ba = 0, bb = 0, ta = 0, tb = 0;
#pragma unroll
    for (int w = 0; w < kTileThreads / 64; ++w) {
        if (w < wave) ba += T.part[0][w], bb += T.part[1][w];
        ta += T.part[0][w], tb += T.part[1][w];
    }
    total = make_uint2(ta, tb);
    return make_uint2(ba + xa - a, bb + xb - b);
}

template <bool FAST>
__global__ __launch_bounds__(kTileThreads) void k_pressure_tile(
    Batch<float> B, uint32_t iter, const uint2* __restrict__ bcells, const float4* __restrict__ bpos,
    const uint2* __restrict__ fcells, const float4* __restrict__ xr, const float4* __restrict__ xp,
    float* __restrict__ dv, float* __restrict__ p, float4* __restrict__ xf) {
    using O = Ops<float, FAST>;
    __shared__ PressureTile T;
    __shared__ unsigned long long lds[2][8];
    const BlockRange rg = B.blocks[blockIdx.x];
    const SphParams<float> P = B.scenes[rg.scene];
    LoopState* st = B.states + rg.scene;
    if (pressure_pass_skipped<float>(st, iter, P)) return;
    const uint2* fc = fcells + P.bucket_base;
    const uint2* bc = bcells + P.bucket_base;
    const uint32_t t = rg.begin + threadIdx.x;
    const bool act = t < rg.end;
    const uint32_t i = act ? t : 0u;
    float4 self = make_float4(0.f, 0.f, 0.f, 0.f), pr = self;
    int3 c = make_int3(0, 0, 0);
    if (act) {
        self = xr[i];
        pr = xp[i];
        c = cell_of4<float>(self, P.cell);
    }
    // 1. the workgroup's buckets
    for (uint32_t q = threadIdx.x; q < kTileSlots; q += kTileThreads) T.key[q] = kTileEmpty;
    if (threadIdx.x == 0) T.overflow = 0u;
    __syncthreads();
    if (act) for_each_bucket(c, P.grid, [&](uint32_t key) { tile_insert(T, key); });
    __syncthreads();
    bool staged = T.overflow == 0u;
    // 2. staged ranges: each thread owns kTileSlots / kTileThreads consecutive slots
    constexpr uint32_t kPer = kTileSlots / kTileThreads;
    uint2 fr[kPer], br[kPer];
    uint32_t fsum = 0, bsum = 0;
    if (staged) {
#pragma unroll
        for (uint32_t q = 0; q < kPer; ++q) {
            const uint32_t key = T.key[threadIdx.x * kPer + q];
            fr[q] = key != kTileEmpty ? fc[key] : make_uint2(0u, 0u);
            br[q] = key != kTileEmpty && P.nb ? bc[key] : make_uint2(0u, 0u);
            fsum += fr[q].y - fr[q].x;
            bsum += br[q].y - br[q].x;
        }
    }
    uint2 total;
    const uint2 base = tile_scan(T, fsum, bsum, total);
    staged = staged && total.x <= kTileFluid && total.y <= kTileBoundary;
    // 3. stage the candidates
    if (staged) {
        uint32_t fo = base.x, bo = base.y;
#pragma unroll
        for (uint32_t q = 0; q < kPer; ++q) {
            const uint32_t fn = fr[q].y - fr[q].x, bn = br[q].y - br[q].x;
            T.fslot[threadIdx.x * kPer + q] = fo | (fn << 16);
            T.bslot[threadIdx.x * kPer + q] = bo | (bn << 16);
            T.fsrc[threadIdx.x * kPer + q] = fr[q].x;
            T.bsrc[threadIdx.x * kPer + q] = br[q].x;
            fo += fn;
            bo += bn;
        }
    }
    __syncthreads();
    if (staged) {
        tile_copy(T.fslot, T.fsrc, T.fx, xp, total.x);
        tile_copy(T.bslot, T.bsrc, T.bx, bpos, total.y);
    }
    __syncthreads();
    unsigned long long m[2] = {0, 0};  // max dv, max expansion
    if (act) {
        float pd = 0.f;
        auto add = [&](const float4& o) { pd = poly6_add<float, FAST>(pd, pr, o, P.h2); };
        if (staged) {
            if (P.nb)
                for_each_bucket(c, P.grid, [&](uint32_t key) {
                    const uint32_t r = T.bslot[tile_find(T, key)];
                    for (uint32_t j = r & 0xFFFFu, e = (r & 0xFFFFu) + (r >> 16); j < e; ++j) add(T.bx[j]);
                });
            for_each_bucket(c, P.grid, [&](uint32_t key) {
                const uint32_t r = T.fslot[tile_find(T, key)];
                for (uint32_t j = r & 0xFFFFu, e = (r & 0xFFFFu) + (r >> 16); j < e; ++j) add(T.fx[j]);
            });
        } else {
            if (P.nb)
                walk_buckets<float>(cell_table(P, bcells, B.bgrid), c, [&](uint32_t j) { return bpos[j]; },
                                    [&](const float4& o, uint32_t) { add(o); });
            walk_buckets<float>(cell_table(P, fcells, B.fgrid), c, [&](uint32_t j) { return xp[j]; },
                                [&](const float4& o, uint32_t) { add(o); });
        }
        pd *= P.density_scale;
        const float v = fmaxf(0.f, pd - P.rho0);
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

// Publication of iteration `iter` by the first workgroup of the pressure-force launch
// (the pressure pass that produced the results has completed): every scene that ran
// the iteration writes its results, the number of scenes that go on is stored, and
// then the step tag, so the host decides how far to enqueue without waiting on the
// stream.  The tag is read from device memory (written at the start of each step),
// so a launch recorded in a graph publishes the tag of the step that replays it.
template <typename R>
__device__ void publish_iteration(const Batch<R>& B, uint32_t iter, HostMirror* host, SceneResult* res,
                                  const unsigned int* tag) {
    __shared__ unsigned int s_run[16];
    unsigned int running = 0;
    for (uint32_t s = threadIdx.x; s < B.scene_count; s += blockDim.x) {
        const SphParams<R>& P = B.scenes[s];
        const LoopState* st = B.states + s;
        if (P.n == 0 || iteration_skipped<R>(st, iter, P)) continue;
        SceneResult r;
        r.max_dv = st->max_dv[iter];
        r.max_exp = st->max_exp[iter];
        r.candidates[0] = st->fluid_candidates;
        r.candidates[1] = st->boundary_candidates;
        r.iterations = iter + 1;
        r.warm_rejected = st->warm_rejected;
        r.first_dv = P.warm ? st->dual_dv[1] : st->max_dv[0];
        r.list_waves = st->list_waves;
        r.running =
            iteration_is_last<R>(st->max_dv, iter, P.min_iterations, P.max_iterations, P.rho0, P.threshold) ? 0u : 1u;
        res[s] = r;
        running += r.running;
    }
    __threadfence_system();
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) running += __shfl_xor(running, off);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) s_run[wave] = running;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int total = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) total += s_run[w];
        host->running[iter] = total;
        __threadfence_system();
        __hip_atomic_store(&host->tag[iter], *tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// update_pressure_force (pcisph.cl:265-300) fused with what follows it: the next
// prediction (pcisph.cl:175-193 into xp, Fluid.cpp:253-260) or, when this iteration
// ends the scene's loop, the final in-place integration (Fluid.cpp:309-316) into
// pos_out/vel_out.  Fluid ranges only.
// One workgroup's fluid range `range` (k_pressure_force, k_correction_loop).
// warm_pass: the pressure force of the network's initial pressures, before the first
// prediction -- not an iteration of the loop.
template <typename R, bool FAST>
__device__ __forceinline__ void pressure_force_range(
    const Batch<R>& B, uint32_t range, uint32_t iter, const uint2* __restrict__ bcells, const V4<R>* __restrict__ bxf,
    const uint2* __restrict__ fcells, const V4<R>* __restrict__ xf, const V4<R>* __restrict__ fo,
    const V4<R>* __restrict__ vel, V4<R>* __restrict__ fp, V4<R>* __restrict__ xp, V4<R>* __restrict__ pos_out,
    V4<R>* __restrict__ vel_out, const R* __restrict__ p, int warm_pass, V4<R>* __restrict__ xp_save,
    const NeighborLists& NL) {
    using O = Ops<R, FAST>;
    const BlockRange rg = B.blocks[range];
    const SphParams<R> P = B.scenes[rg.scene];
    LoopState* st = B.states + rg.scene;
    if (warm_pass ? !P.warm : iteration_skipped<R>(st, iter, P)) return;
    const bool last = !warm_pass && iteration_is_last<R>(st->max_dv, iter, P.min_iterations, P.max_iterations,
                                                         P.rho0, P.threshold);
    const uint32_t t = rg.begin + threadIdx.x;
    float moved = 0.f;
    uint32_t bucket = 0xFFFFFFFFu;
    if (t < rg.end) {
        const uint32_t i = t;
        const V4<R> self = xf[i];
        const R sf = self.w;
        const int3 c = cell_of4<R>(self, P.cell);
        bucket = P.bucket_base + hash_key((uint32_t)c.x, (uint32_t)c.y, (uint32_t)c.z, P.grid);
        R fx = 0, fy = 0, fz = 0;
        auto add = [&](const V4<R>& o) {
            R dx, dy, dz;
            const R r2 = dist2<R>(self.x, self.y, self.z, o.x, o.y, o.z, dx, dy, dz);
            if (r2 > P.h2) return;
            if constexpr (FAST) {
                // one reciprocal square root for r and 1/r: the same terms to within the
                // fast mode's rounding (its 0.0001 cut taken on r^2)
                if (r2 < 1e-8f) return;  // kernel_spiky_d1 (pcisph.cl:28-37)
                const float q = __builtin_amdgcn_rsqf(r2);
                const float r = r2 * q;
                if (r > P.h) return;
                const float dr = P.h - r;
                const float sc = dr * dr * q * (sf + o.w);
                fx += dx * sc;
                fy += dy * sc;
                fz += dz * sc;
            } else {
                const R r = O::sqrt(r2);
                if (r > P.h || r < R(0.0001f)) return;  // kernel_spiky_d1 (pcisph.cl:28-37)
                const R dr = P.h - r;
                const R factor = sf + o.w;
                fx += O::div(dx * dr * dr, r) * factor;
                fy += O::div(dy * dr * dr, r) * factor;
                fz += O::div(dz * dr * dr, r) * factor;
            }
        };
        // positions here are x, where the lists are exact; their chunks stream past L2
        // (non-temporal), which keeps this pass's writes of x^ there for the next
        // pressure pass (measured: 1M transient pressure pass -4%)
        const bool lists = NL.on != 0;
        if (P.nb) {
            const uint32_t len = lists ? NL.bcount[t] : kListOverflow;
            walk<R, kListBatch, true>(NL.b, len, B.n_total, t, cell_table(P, bcells, B.bgrid), c,
                    [&](uint32_t j) { return bxf[j]; }, [&](const V4<R>& o, uint32_t) { add(o); });
        }
        const uint32_t len = lists ? NL.fcount[t] : kListOverflow;
        walk<R, kListBatch, true>(NL.f, len, B.n_total, t, cell_table(P, fcells, B.fgrid), c,
                                  [&](uint32_t j) { return xf[j]; },
                [&](const V4<R>& o, uint32_t) { add(o); });
        V4<R> f;
        f.x = fx * P.pf_scale;
        f.y = fy * P.pf_scale;
        f.z = fz * P.pf_scale;
        f.w = 0;
        const V4<R> o = fo[i];
        const V4<R> v = vel[i];
        V4<R> xo, vo;
        integrate<R, FAST>(P, self, v, o.x + f.x, o.y + f.y, o.z + f.z, xo, vo);
        if (!last) {
            if (warm_pass) xp_save[i] = xp[i];  // keep the plain start's prediction
            xp[i] = xo;
            moved = displacement<R>(xo, self);
        } else {
            fp[i] = f;  // the step's pressure force: only the last pass's is ever read
            pos_out[i] = xo;
            vo.w = p[i];  // the converged pressure rides along to the next step
            vel_out[i] = vo;
            const bool bad = !(isfinite(xo.x) && isfinite(xo.y) && isfinite(xo.z) && isfinite(vo.x) &&
                               isfinite(vo.y) && isfinite(vo.z));
            if (bad) atomicOr(&st->nonfinite, 1u);
        }
    }
    if (!last && !warm_pass && NL.on && P.disp_limit > 0.f)
        record_displacement(NL.disp[(iter + 1) & 1], bucket, moved);
}

template <typename R, bool FAST>
__global__ __launch_bounds__(kMaxBlockThreads) void k_pressure_force(Batch<R> B, uint32_t iter, const uint2* __restrict__ bcells,
                                 const V4<R>* __restrict__ bxf, const uint2* __restrict__ fcells,
                                 const V4<R>* __restrict__ xf, const V4<R>* __restrict__ fo,
                                 const V4<R>* __restrict__ vel, V4<R>* __restrict__ fp, V4<R>* __restrict__ xp,
                                 V4<R>* __restrict__ pos_out, V4<R>* __restrict__ vel_out,
                                 const R* __restrict__ p, HostMirror* host, SceneResult* res,
                                 const unsigned int* tag, int warm_pass, V4<R>* __restrict__ xp_save,
                                 NeighborLists NL) {
    if (blockIdx.x == 0 && !warm_pass) publish_iteration<R>(B, iter, host, res, tag);
    pressure_force_range<R, FAST>(B, blockIdx.x, iter, bcells, bxf, fcells, xf, fo, vel, fp, xp, pos_out, vel_out, p,
                                  warm_pass, xp_save, NL);
}

// The network's start: p = max(0, p_prev + y p_ref), y the inference kernel's output
// (held in p) and p_prev the particle's converged pressure of the previous step
// (the velocity's w); then the force pass's p/rho^2.  A network that outputs 0
// starts every particle from its previous pressure.  Fluid ranges of the scenes
// the step starts from the network.
template <typename R>
__global__ void k_warm_apply(Batch<R> B, const V4<R>* __restrict__ xr, const V4<R>* __restrict__ vel,
                             R* __restrict__ p, V4<R>* __restrict__ xf) {
    const BlockRange rg = B.blocks[blockIdx.x];
    const uint32_t i = rg.begin + threadIdx.x;
    if (i >= rg.end) return;
    const SphParams<R>& P = B.scenes[rg.scene];
    if (!P.warm) return;  // a scene left out of the warm start keeps the plain start
    const R pi = fmax(R(0), vel[i].w + (R)((float)p[i] * P.p_ref));
    p[i] = pi;
    V4<R> o = xr[i];
    o.w = pi / (o.w * o.w);
    xf[i] = o;
}

// First iteration of a warm-started step: the pressure update of pcisph.cl:212-263
// for two starts at once -- the network's (xp, p) and the plain one (xp0, p = 0) --
// sharing the candidate loop.  Each scene continues with whichever start has the
// smaller residual max |rho* - rho0| (k_select_start), so a prediction that makes
// the residual worse is dropped and the loop is exactly the plain one.
// The dual pass's results of fluid particle i from its two sums (sw from the
// network's start, sz from the plain start): density variations, pressures and
// p/rho^2 of both, and the maxima m (dv warm, expansion warm, dv plain, expansion plain).
template <typename R, bool FAST>
__device__ __forceinline__ void dual_fluid_outputs(const SphParams<R>& P, uint32_t i, const V4<R>& self, R sw, R sz,
                                                   R* __restrict__ dv, R* __restrict__ p, V4<R>* __restrict__ xf,
                                                   const DualArrays<R>& D, unsigned long long (&m)[4]) {
    using O = Ops<R, FAST>;
    sw *= P.density_scale;
    sz *= P.density_scale;
    const R vw = fmax(R(0), sw - P.rho0), vz = fmax(R(0), sz - P.rho0);
    m[0] = Bits<R>::of(vw);
    const bool bulk = self.w >= P.rho0;  // as in k_pressure: no free-surface particles
    m[1] = bulk ? Bits<R>::of(fmax(R(0), P.rho0 - sw)) : 0ull;
    m[2] = Bits<R>::of(vz);
    m[3] = bulk ? Bits<R>::of(fmax(R(0), P.rho0 - sz)) : 0ull;
    dv[i] = vw;
    D.dv0[i] = vz;
    R pi = p[i];
    if (vw != R(0)) {
        pi += vw * P.delta;
        p[i] = pi;
    }
    const R pz0 = vz != R(0) ? R(0) + vz * P.delta : R(0);
    D.p0[i] = pz0;
    V4<R> out = self;
    out.w = O::div(pi, self.w * self.w);
    xf[i] = out;
    out.w = O::div(pz0, self.w * self.w);
    D.xf0[i] = out;
}

// The dual pass of boundary particle b (boundary mode, pcisph.cl:231-236, for both
// starts).
template <typename R, bool FAST>
__device__ __forceinline__ void dual_boundary_particle(const SphParams<R>& P, uint32_t b, const CellTable& FT,
                                                       const V4<R>* __restrict__ bpos, const R* __restrict__ binit,
                                                       const V4<R>* __restrict__ xp, R* __restrict__ bp,
                                                       V4<R>* __restrict__ bxf, const DualArrays<R>& D) {
    using O = Ops<R, FAST>;
    const V4<R> self = bpos[b];
    const int3 c = cell_of4<R>(self, P.cell);
    R sw = 0, sz = 0;
    sw += binit[b];
    sz += binit[b];
    struct Pair {
        V4<R> w, z;
    };
    walk_buckets<R>(FT, c, [&](uint32_t j) { return Pair{xp[j], D.xp0[j]}; },
                    [&](const Pair& o, uint32_t) {
                        sw = poly6_add<R, FAST>(sw, self, o.w, P.h2);
                        sz = poly6_add<R, FAST>(sz, self, o.z, P.h2);
                    });
    sw *= P.density_scale;
    sz *= P.density_scale;
    const R vw = fmax(R(0), sw - P.rho0), vz = fmax(R(0), sz - P.rho0);
    const R pbw = vw != R(0) ? bp[b] + vw * P.delta : bp[b];
    const R pbz = vz != R(0) ? R(0) + vz * P.delta : R(0);
    bp[b] = pbw;
    D.bp0[b] = pbz;
    V4<R> out = self;
    out.w = O::div(pbw, P.rho0_sq);
    bxf[b] = out;
    out.w = O::div(pbz, P.rho0_sq);
    D.bxf0[b] = out;
}

template <typename R, bool FAST>
__global__ __launch_bounds__(kMaxBlockThreads) void k_pressure_dual(Batch<R> B, const uint2* __restrict__ bcells, const V4<R>* __restrict__ bpos,
                                const R* __restrict__ binit, const uint2* __restrict__ fcells,
                                const V4<R>* __restrict__ xr, const V4<R>* __restrict__ xp, R* __restrict__ dv,
                                R* __restrict__ p, V4<R>* __restrict__ xf, R* __restrict__ bp,
                                V4<R>* __restrict__ bxf, DualArrays<R> D) {
    using O = Ops<R, FAST>;
    __shared__ unsigned long long lds[4][8];
    if (D.disp_next) clear_table(D.disp_next, D.buckets);  // for iteration 1's prediction
    const BlockRange rg = B.blocks[blockIdx.x];
    const SphParams<R> P = B.scenes[rg.scene];
    if (!P.warm) return;  // scenes left out of the warm start: the plain pass runs iteration 0
    LoopState* st = B.states + rg.scene;
    const CellTable FT = cell_table(P, fcells, B.fgrid);
    const uint32_t t = rg.begin + threadIdx.x;
    const bool boundary = (rg.flags & kRangeBoundary) != 0;
    const uint32_t i = t;
    unsigned long long m[4] = {0, 0, 0, 0};  // dv warm, exp warm, dv plain, exp plain
    if (!boundary) {
        if (t < rg.end) {
            const V4<R> self = xr[i];
            const V4<R> pw = xp[i], pz = D.xp0[i];
            const int3 c = cell_of4<R>(self, P.cell);
            R sw = 0, sz = 0;
            if (P.nb) {
                CellTable BT = cell_table(P, bcells, B.bgrid);
                BT.prune = far_prunable<R>(displacement<R>(pw, self), (float)P.h, P.cell) &&
                           far_prunable<R>(displacement<R>(pz, self), (float)P.h, P.cell);
                if (!(BT.prune && quiet_cell(BT, c))) {
                    walk_buckets<R>(BT, c, [&](uint32_t j) { return bpos[j]; },
                                    [&](const V4<R>& o, uint32_t) {
                                        sw = poly6_add<R, FAST>(sw, pw, o, P.h2);
                                        sz = poly6_add<R, FAST>(sz, pz, o, P.h2);
                                    });
                }
            }
            struct Pair {
                V4<R> w, z;
            };
            walk_buckets<R>(FT, c, [&](uint32_t j) { return Pair{xp[j], D.xp0[j]}; },
                            [&](const Pair& o, uint32_t) {
                                sw = poly6_add<R, FAST>(sw, pw, o.w, P.h2);
                                sz = poly6_add<R, FAST>(sz, pz, o.z, P.h2);
                            });
            dual_fluid_outputs<R, FAST>(P, i, self, sw, sz, dv, p, xf, D, m);
        }
    } else if (i < rg.end) {
        dual_boundary_particle<R, FAST>(P, i, FT, bpos, binit, xp, bp, bxf, D);
    }
    // 0: dv warm, 1: expansion warm, 2: dv plain, 3: expansion plain
    unsigned long long* const dst[4] = {&st->dual_dv[0], &st->dual_exp[0], &st->dual_dv[1], &st->dual_exp[1]};
    block_max_to<4>(m, dst, lds);
}

// Per warm-started scene: picks the start with the smaller residual and records its
// iteration-0 results where the loop's termination logic reads them.
template <typename R>
__global__ void k_select_start(const SphParams<R>* __restrict__ scenes, LoopState* states, uint32_t scene_count) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < scene_count; s += gridDim.x * blockDim.x) {
        if (!scenes[s].warm) continue;  // iteration 0 was the plain pass's
        LoopState* st = states + s;
        const R dw = Bits<R>::to(st->dual_dv[0]), ew = Bits<R>::to(st->dual_exp[0]);
        const R dz = Bits<R>::to(st->dual_dv[1]), ez = Bits<R>::to(st->dual_exp[1]);
        const R rw = dw > ew ? dw : ew, rz = dz > ez ? dz : ez;
        const bool reject = !(rw < rz);
        st->warm_rejected = reject ? 1u : 0u;
        st->max_dv[0] = reject ? st->dual_dv[1] : st->dual_dv[0];
        st->max_exp[0] = reject ? st->dual_exp[1] : st->dual_exp[0];
    }
}

// After k_select_start rejected the network's start of a scene: make the plain
// start's arrays the live ones (fluid and boundary ranges).
template <typename R>
__global__ void k_adopt_plain_start(Batch<R> B, DualArrays<R> D, V4<R>* __restrict__ xp, V4<R>* __restrict__ xf,
                                    R* __restrict__ p, R* __restrict__ dv, R* __restrict__ bp,
                                    V4<R>* __restrict__ bxf) {
    const BlockRange rg = B.blocks[blockIdx.x];
    if (!B.states[rg.scene].warm_rejected) return;
    const uint32_t i = rg.begin + threadIdx.x;
    if (i >= rg.end) return;
    if (!(rg.flags & kRangeBoundary)) {
        xp[i] = D.xp0[i];
        xf[i] = D.xf0[i];
        p[i] = D.p0[i];
        dv[i] = D.dv0[i];
    } else {
        bp[i] = D.bp0[i];
        bxf[i] = D.bxf0[i];
    }
}

// initialize_boundary_boundary_pred_densities (pcisph.cl:195-210), opt-in; `blocks`
// is the boundary part of the block table.
template <typename R, bool FAST>
__global__ __launch_bounds__(kMaxBlockThreads) void k_boundary_init(const SphParams<R>* __restrict__ scenes, const BlockRange* __restrict__ blocks,
                                const uint2* __restrict__ bcells, const V4<R>* __restrict__ bpos,
                                R* __restrict__ binit) {
    const BlockRange rg = blocks[blockIdx.x];
    const uint32_t b = rg.begin + threadIdx.x;
    if (b >= rg.end) return;
    const SphParams<R> P = scenes[rg.scene];
    const V4<R> self = bpos[b];
    const int3 c = cell_of4<R>(self, P.cell);
    R r = 0;
    for_each_candidate<R>(bcells + P.bucket_base, c, P.grid, [&](uint32_t j) {
        const V4<R> o = bpos[j];
        R dx, dy, dz;
        r += poly6<R>(dist2<R>(self.x, self.y, self.z, o.x, o.y, o.z, dx, dy, dz), P.h2);
    });
    binit[b] = r;
}

}  // namespace pcisph
