// sph_launch_impl.h - definitions of the launchers declared in sph_launch.h.
// Included once per arithmetic-mode translation unit (sph_ieee.hip, sph_fast.hip).
#pragma once

#include <type_traits>

#include "sph_launch.h"

namespace pcisph {

inline uint32_t grid_for(uint32_t n, uint32_t block) { return (n + block - 1) / block; }

template <typename R, bool FAST>
void SphLaunch<R, FAST>::density(const Launch& L, const DeviceArrays<R>& A) {
    const Batch<R>& B = A.batch;
    const uint32_t grid = B.fluid_blocks + B.boundary_blocks;
    if (!grid) return;
    if constexpr (std::is_same<R, float>::value) {
        if (A.staged_pressure) {
            // fluid ranges from the step's LDS staging (set up before this pass)
            hipLaunchKernelGGL((k_density_staged<FAST>), dim3(grid), dim3(L.block), 0, L.stream, B, A.bcells,
                               A.bpos, A.cells, A.pos, A.xr, A.counts, A.bp, A.bxf, A.NL, A.stage);
            return;
        }
    }
    hipLaunchKernelGGL((k_density<R, FAST>), dim3(grid), dim3(L.block), 0, L.stream, B, A.bcells, A.bpos, A.cells,
                       A.pos, A.xr, A.counts, A.bp, A.bxf, A.NL);
}

template <typename R, bool FAST>
void SphLaunch<R, FAST>::normal(const Launch& L, const DeviceArrays<R>& A) {
    const Batch<R>& B = A.batch;
    if (!B.fluid_blocks) return;
    hipLaunchKernelGGL((k_normal<R, FAST>), dim3(B.fluid_blocks), dim3(L.block), 0, L.stream, B, A.cells, A.xr,
                       A.nrm, A.NL);
}

template <typename R, bool FAST>
void SphLaunch<R, FAST>::force_init(const Launch& L, const DeviceArrays<R>& A, const FeatureArgs& F) {
    const Batch<R>& B = A.batch;
    if (!B.fluid_blocks) return;
    hipLaunchKernelGGL((k_force_init<R, FAST>), dim3(B.fluid_blocks), dim3(L.block), 0, L.stream, B, A.cells, A.xr,
                       A.vel, A.nrm, A.fo, A.p, A.xp, F, A.NL);
}

template <typename R, bool FAST>
void SphLaunch<R, FAST>::pressure(const Launch& L, const DeviceArrays<R>& A, uint32_t iter) {
    const Batch<R>& B = A.batch;
    const uint32_t grid = B.fluid_blocks + B.boundary_blocks;
    if (!grid) return;
    if constexpr (std::is_same<R, float>::value) {
        if (A.staged_pressure) {
            // fluid ranges from the step's LDS staging, boundary ranges in the same launch
            hipLaunchKernelGGL((k_pressure_staged<FAST>), dim3(grid), dim3(L.block), 0, L.stream, B, iter, A.bcells,
                               A.bpos, A.binit, A.cells, A.xr, A.xp, A.dv, A.p, A.xf, A.bp, A.bxf, A.stage, A.NL);
            return;
        }
        if (A.tile_pressure && L.block == (uint32_t)kTileThreads) {
            // fluid ranges from LDS tiles, boundary ranges as below
            if (B.fluid_blocks)
                hipLaunchKernelGGL((k_pressure_tile<FAST>), dim3(B.fluid_blocks), dim3(kTileThreads), 0, L.stream,
                                   B, iter, A.bcells, A.bpos, A.cells, A.xr, A.xp, A.dv, A.p, A.xf);
            if (B.boundary_blocks) {
                Batch<R> Bb = B;
                Bb.blocks = B.blocks + B.fluid_blocks;
                Bb.fluid_blocks = 0;
                hipLaunchKernelGGL((k_pressure<R, FAST>), dim3(B.boundary_blocks), dim3(L.block), 0, L.stream, Bb,
                                   iter, A.bcells, A.bpos, A.binit, A.cells, A.xr, A.xp, A.dv, A.p, A.xf, A.bp,
                                   A.bxf, A.NL);
            }
            return;
        }
    }
    hipLaunchKernelGGL((k_pressure<R, FAST>), dim3(grid), dim3(L.block), 0, L.stream, B, iter, A.bcells, A.bpos,
                       A.binit, A.cells, A.xr, A.xp, A.dv, A.p, A.xf, A.bp, A.bxf, A.NL);
}

template <typename R, bool FAST>
void SphLaunch<R, FAST>::pressure_force(const Launch& L, const DeviceArrays<R>& A, uint32_t iter, HostMirror* host,
                                        SceneResult* res, const unsigned int* tag, int warm_pass,
                                        V4<R>* xp_save) {
    const Batch<R>& B = A.batch;
    if (!B.fluid_blocks) return;
    hipLaunchKernelGGL((k_pressure_force<R, FAST>), dim3(B.fluid_blocks), dim3(L.block), 0, L.stream, B, iter,
                       A.bcells, A.bxf, A.cells, A.xf, A.fo, A.vel, A.fp, A.xp, A.pos_in, A.vel_in, A.p, host, res,
                       tag, warm_pass, xp_save, A.NL);
}

template <typename R, bool FAST>
hipError_t SphLaunch<R, FAST>::correction_loop(const Launch& L, const DeviceArrays<R>& A, uint32_t maxit,
                                               HostMirror* host, SceneResult* res, const unsigned int* tag,
                                               const GridSync& g) {
    Batch<R> B = A.batch;
    const uint32_t total = B.fluid_blocks + B.boundary_blocks;
    if (!total || !maxit) return hipSuccess;
    LoopArgs<R> a{A.bcells, A.bpos, A.binit, A.cells, A.xr, A.xp, A.dv, A.p, A.xf, A.bp, A.bxf, A.fo, A.vel, A.fp,
                  A.pos_in, A.vel_in, host, res, tag, A.NL, A.stage};
    GridSync sync = g;
    const void* kernel = (const void*)k_correction_loop<R, FAST, false>;
    if constexpr (std::is_same<R, float>::value)
        if (A.staged_pressure) kernel = (const void*)k_correction_loop<float, FAST, true>;
    int device = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, (int)L.block, 0);
    if (e != hipSuccess) return e;
    const uint64_t resident = (uint64_t)(per_cu > 0 ? per_cu : 0) * (uint64_t)(cus > 0 ? cus : 0);
    if (!resident) return hipErrorCooperativeLaunchTooLarge;
    const uint32_t grid = (uint32_t)(total < resident ? total : resident);
    void* args[] = {(void*)&B, (void*)&maxit, (void*)&a, (void*)&sync};
    return hipLaunchCooperativeKernel(kernel, dim3(grid), dim3(L.block), args, 0, L.stream);
}

template <typename R, bool FAST>
void SphLaunch<R, FAST>::pressure_dual(const Launch& L, const DeviceArrays<R>& A, const DualArrays<R>& D) {
    const Batch<R>& B = A.batch;
    const uint32_t grid = B.fluid_blocks + B.boundary_blocks;
    if (!grid) return;
    bool done = false;
    if constexpr (std::is_same<R, float>::value) {
        if (A.staged_pressure) {
            // fluid ranges from the step's LDS staging
            hipLaunchKernelGGL((k_pressure_dual_staged<FAST>), dim3(grid), dim3(L.block), 0, L.stream, B, A.bcells,
                               A.bpos, A.binit, A.cells, A.xr, A.xp, A.dv, A.p, A.xf, A.bp, A.bxf, D, A.stage);
            done = true;
        }
    }
    if (!done)
        hipLaunchKernelGGL((k_pressure_dual<R, FAST>), dim3(grid), dim3(L.block), 0, L.stream, B, A.bcells, A.bpos,
                           A.binit, A.cells, A.xr, A.xp, A.dv, A.p, A.xf, A.bp, A.bxf, D);
    hipLaunchKernelGGL((k_select_start<R>), dim3(grid_for(B.scene_count, 256)), dim3(256), 0, L.stream, B.scenes,
                       B.states, B.scene_count);
    hipLaunchKernelGGL((k_adopt_plain_start<R>), dim3(grid), dim3(L.block), 0, L.stream, B, D, A.xp, A.xf, A.p,
                       A.dv, A.bp, A.bxf);
}

template <typename R, bool FAST>
void SphLaunch<R, FAST>::boundary_init(const Launch& L, const DeviceArrays<R>& A) {
    const Batch<R>& B = A.batch;
    if (!B.boundary_blocks) return;
    hipLaunchKernelGGL((k_boundary_init<R, FAST>), dim3(B.boundary_blocks), dim3(L.block), 0, L.stream, B.scenes,
                       B.blocks + B.fluid_blocks, A.bcells, A.bpos, A.binit);
}

}  // namespace pcisph
