// sph_persistent.h - the pressure-correction loop (PCISPH/src/sim/Fluid.cpp:249-307)
// as one cooperative launch.
//
// Each iteration is the two passes of the launch-per-pass loop -- update_pressure
// (pcisph.cl:212-263) over every range of the block table, then update_pressure_force
// with the next prediction or the final integration over the fluid ranges -- run by
// the same per-range code (pressure_range / pressure_staged_range /
// pressure_force_range), each workgroup taking ranges blockIdx.x, blockIdx.x +
// gridDim.x, ...  A grid-wide barrier separates the passes; after the second one every
// workgroup reads the same termination state (the scenes' per-iteration maxima,
// complete once every pressure pass has arrived) and leaves the loop together.  The
// first workgroup publishes each iteration as the pressure-force launch does, so the
// host follows the loop the same way.  Results are those of the launch-per-pass loop
// bit for bit: the same code on the same data in the same order.
//
// The launch is cooperative (every workgroup resident, or the launch fails), so the
// barrier cannot wait on a workgroup that never runs; a wait longer than the given
// wall-clock bound still ends it, sets the abort word and makes every workgroup leave,
// and the host reports the failure.
#pragma once

#include "sph_staged.h"

namespace pcisph {

// Barrier words in device memory, cleared before each launch: [0] arrivals,
// [1] generation, [2] abort.
struct GridSync {
    unsigned int* words;
    unsigned long long timeout;  // wall_clock64() ticks one wait may take
};

// All workgroups of the launch meet here; false after an abort.  Thread 0 of each
// workgroup arrives (releasing the workgroup's writes at device scope) and waits for
// the generation to move on (acquiring the other workgroups' writes).
__device__ __forceinline__ bool grid_barrier(const GridSync& g) {
    __shared__ unsigned int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int* const count = g.words;
        unsigned int* const gen = g.words + 1;
        unsigned int* const abort = g.words + 2;
        const unsigned int now = __hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        const unsigned int arrived = __hip_atomic_fetch_add(count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        bool ok = true;
        if (arrived == gridDim.x - 1) {
            __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(gen, now + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const unsigned long long t0 = (unsigned long long)wall_clock64();
            while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == now) {
                if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    ok = false;
                    break;
                }
                if ((unsigned long long)wall_clock64() - t0 > g.timeout) {
                    __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __threadfence();
        s_ok = ok ? 1u : 0u;
    }
    __syncthreads();
    return s_ok != 0u;
}

// Everything one iteration reads and writes (the arguments of the two passes).
template <typename R>
struct LoopArgs {
    const uint2* bcells;
    const V4<R>* bpos;
    const R* binit;
    const uint2* fcells;
    const V4<R>* xr;
    V4<R>* xp;
    R* dv;
    R* p;
    V4<R>* xf;
    R* bp;
    V4<R>* bxf;
    const V4<R>* fo;
    const V4<R>* vel;
    V4<R>* fp;
    V4<R>* pos_out;
    V4<R>* vel_out;
    HostMirror* host;
    SceneResult* res;
    const unsigned int* tag;
    NeighborLists NL;
    StageMaps S;
};

// The correction loop, iterations 0 .. maxit - 1 until every scene's loop has ended.
// STAGED: the FP32 pressure pass from the step's LDS staging (k_pressure_staged),
// else k_pressure's.
template <typename R, bool FAST, bool STAGED>
__global__ __launch_bounds__(kMaxBlockThreads) void k_correction_loop(Batch<R> B, uint32_t maxit, LoopArgs<R> a,
                                                                      GridSync g) {
    const uint32_t total = B.fluid_blocks + B.boundary_blocks;
    for (uint32_t it = 0; it < maxit; ++it) {
        if (a.NL.on) clear_table(a.NL.disp[(it + 1) & 1], a.NL.buckets);  // for the next prediction
        for (uint32_t r = blockIdx.x; r < total; r += gridDim.x) {
            if constexpr (STAGED)
                pressure_staged_range<FAST>(B, r, it, a.bcells, a.bpos, a.binit, a.fcells, a.xr, a.xp, a.dv, a.p,
                                            a.xf, a.bp, a.bxf, a.S);
            else
                pressure_range<R, FAST>(B, r, it, a.bcells, a.bpos, a.binit, a.fcells, a.xr, a.xp, a.dv, a.p, a.xf,
                                        a.bp, a.bxf, a.NL);
            __syncthreads();
        }
        if (!grid_barrier(g)) return;
        if (blockIdx.x == 0) publish_iteration<R>(B, it, a.host, a.res, a.tag);
        for (uint32_t r = blockIdx.x; r < B.fluid_blocks; r += gridDim.x) {
            pressure_force_range<R, FAST>(B, r, it, a.bcells, a.bxf, a.fcells, a.xf, a.fo, a.vel, a.fp, a.xp,
                                          a.pos_out, a.vel_out, a.p, 0, nullptr, a.NL);
            __syncthreads();
        }
        if (!grid_barrier(g)) return;
        bool running = false;
        for (uint32_t s = 0; s < B.scene_count && !running; ++s) {
            const SphParams<R>& P = B.scenes[s];
            running = P.n && !iteration_is_last<R>(B.states[s].max_dv, it, P.min_iterations, P.max_iterations,
                                                   P.rho0, P.threshold);
        }
        if (!running) return;
    }
}

}  // namespace pcisph
