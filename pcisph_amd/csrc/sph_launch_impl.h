// sph_launch_impl.h - definitions of the launchers declared in sph_launch.h.
// Included once per arithmetic-mode translation unit (sph_ieee.hip, sph_fast.hip).
#pragma once

#include "sph_launch.h"

namespace pcisph {

inline uint32_t grid_for(uint32_t n, uint32_t block) { return (n + block - 1) / block; }

template <typename R, bool FAST>
void SphLaunch<R, FAST>::density(const Launch& L, const SphParams<R>& P, const DeviceArrays<R>& A) {
    const uint32_t total = P.n + P.nb;
    hipLaunchKernelGGL((k_density<R, FAST>), dim3(grid_for(total, L.block)), dim3(L.block), 0, L.stream, P,
                       A.bcells, A.bpos, A.cells, A.pos, A.xr, A.counts, A.bp, A.bxf, A.state, A.NL);
}

template <typename R, bool FAST>
void SphLaunch<R, FAST>::normal(const Launch& L, const SphParams<R>& P, const DeviceArrays<R>& A) {
    hipLaunchKernelGGL((k_normal<R, FAST>), dim3(grid_for(P.n, L.block)), dim3(L.block), 0, L.stream, P,
                       A.cells, A.xr, A.nrm, A.NL);
}

template <typename R, bool FAST>
void SphLaunch<R, FAST>::force_init(const Launch& L, const SphParams<R>& P, const DeviceArrays<R>& A,
                                    const FeatureArgs& F) {
    hipLaunchKernelGGL((k_force_init<R, FAST>), dim3(grid_for(P.n, L.block)), dim3(L.block), 0, L.stream,
                       P, A.cells, A.xr, A.vel, A.nrm, A.fo, A.p, A.fp, A.xp, F, A.NL, A.state);
}

template <typename R, bool FAST>
void SphLaunch<R, FAST>::pressure(const Launch& L, const SphParams<R>& P, const DeviceArrays<R>& A,
                                  uint32_t iter) {
    const uint32_t total = P.n + P.nb;
    hipLaunchKernelGGL((k_pressure<R, FAST>), dim3(grid_for(total, L.block)), dim3(L.block), 0, L.stream, P,
                       iter, A.state, A.bcells, A.bpos, A.binit, A.cells, A.xr, A.xp, A.dv, A.p, A.xf, A.bp,
                       A.bxf, A.NL);
}

template <typename R, bool FAST>
void SphLaunch<R, FAST>::pressure_force(const Launch& L, const SphParams<R>& P, const DeviceArrays<R>& A,
                                        uint32_t iter, HostMirror* host, unsigned int tag, int warm_pass,
                                        V4<R>* xp_save) {
    hipLaunchKernelGGL((k_pressure_force<R, FAST>), dim3(grid_for(P.n, L.block)), dim3(L.block), 0,
                       L.stream, P, iter, A.state, A.bcells, A.bxf, A.cells, A.xf, A.fo, A.vel, A.fp, A.xp,
                       A.pos_in, A.vel_in, A.p, host, tag, warm_pass, xp_save, A.NL);
}

template <typename R, bool FAST>
void SphLaunch<R, FAST>::pressure_dual(const Launch& L, const SphParams<R>& P, const DeviceArrays<R>& A,
                                       const DualArrays<R>& D) {
    const uint32_t total = P.n + P.nb;
    hipLaunchKernelGGL((k_pressure_dual<R, FAST>), dim3(grid_for(total, L.block)), dim3(L.block), 0, L.stream,
                       P, A.state, A.bcells, A.bpos, A.binit, A.cells, A.xr, A.xp, A.dv, A.p, A.xf, A.bp, A.bxf,
                       D);
    hipLaunchKernelGGL((k_select_start<R>), dim3(1), dim3(64), 0, L.stream, A.state);
    hipLaunchKernelGGL((k_adopt_plain_start<R>), dim3(grid_for(total, 256)), dim3(256), 0, L.stream, P.n, P.nb,
                       A.state, D, A.xp, A.xf, A.p, A.dv, A.bp, A.bxf);
}

template <typename R, bool FAST>
void SphLaunch<R, FAST>::boundary_init(const Launch& L, const SphParams<R>& P, const DeviceArrays<R>& A) {
    hipLaunchKernelGGL((k_boundary_init<R, FAST>), dim3(grid_for(P.nb, L.block)), dim3(L.block), 0, L.stream,
                       P, A.bcells, A.bpos, A.binit);
}

}  // namespace pcisph
