/*
 * oracle.h - C ABI of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * The oracle is a plain C++ restatement of the reference's simulation step
 * (PCISPH/src/sim/Fluid.cpp:108-466 and the kernels in
 * PCISPH/data/kernels/{grid_utils,sort_utils,pcisph}.cl) used only to check the
 * HIP path: by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg.  Nothing in the product links or calls it.
 *
 * The entry points mirror include/pcisph.h one for one (prefix `oracle_`), so a
 * test drives both through the same sequence of calls.
 */
#ifndef PCISPH_ORACLE_H
#define PCISPH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/pcisph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_sim oracle_sim;

/* fp64 = 0: FP32 instantiation in the reference's order of operations;
 * fp64 = 1: the same code on double (positions, velocities, sums). */
oracle_sim* oracle_create(int fp64);
void oracle_destroy(oracle_sim* s);
int oracle_set_threads(int n); /* OpenMP threads used by the per-particle loops */

int oracle_set_boundary_count(oracle_sim* s, uint32_t v);
int oracle_set_fluid_count(oracle_sim* s, uint32_t v);
int oracle_set_delta_t(oracle_sim* s, float v);
int oracle_set_rest_density(oracle_sim* s, float v);
int oracle_set_particle_radius(oracle_sim* s, float v);
int oracle_set_gravity(oracle_sim* s, float v);
int oracle_set_viscosity(oracle_sim* s, float v);
int oracle_set_surface_tension(oracle_sim* s, float v);
int oracle_set_density_variation_threshold(oracle_sim* s, float v);
int oracle_get_params(const oracle_sim* s, Simulation_Params* out);
int oracle_set_solver_options(oracle_sim* s, const pcisph_solver_options* opt);

int oracle_field_size(const oracle_sim* s, int field, size_t* out);
int oracle_upload(oracle_sim* s, int field, const void* host, size_t count);
int oracle_download(oracle_sim* s, int field, void* host, size_t count);
/* FP64 view of a float field (exact widening for the FP32 instantiation). */
int oracle_download_f64(oracle_sim* s, int field, double* host, size_t count);
int oracle_step(oracle_sim* s, pcisph_step_stats* stats);

/* Pure functions, for unit tests of the grid. */
uint32_t oracle_hash_key(int32_t cx, int32_t cy, int32_t cz, uint32_t bucket_count);
void oracle_cell_pos(float x, float y, float z, float cell_size, int32_t out[3]);
uint32_t oracle_sort_bit_count(uint32_t elements);
int oracle_compute_params(float particle_radius, float rest_density, uint32_t boundary_count,
                          uint32_t fluid_count, float viscosity, float surface_tension,
                          float gravity, float delta_t, Simulation_Params* out);

#ifdef __cplusplus
}
#endif

#endif
