/*
 * pcisph_params.h - the 80-byte simulation parameter block shared by the host,
 * the HIP kernels and the CPU oracle.
 *
 * Byte-for-byte the layout of the reference's `Simulation_Params`
 * (PCISPH/data/kernels/Simulation_Params.h:17-45): twenty packed 4-byte fields,
 * float unless marked uint32.  `pcisph_get_params` returns exactly these bytes,
 * so a caller that memcpy's the reference struct can read ours unchanged.
 */
#ifndef PCISPH_PARAMS_H
#define PCISPH_PARAMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#pragma pack(push, 1)
typedef struct Simulation_Params {
    float particle_radius;                 /*  0 */
    float particle_mass;                   /*  4 */
    float rest_density;                    /*  8 */
    uint32_t boundary_count;               /* 12 */
    uint32_t fluid_count;                  /* 16 */
    float kernel_radius;                   /* 20 */
    float kernel_radius2;                  /* 24 */
    uint32_t bucket_count;                 /* 28 */
    float cell_size;                       /* 32 */
    float poly6_normalization;             /* 36 */
    float poly6_d1_normalization;          /* 40 */
    float viscosity_d2_normalization;      /* 44 */
    float spiky_d1_normalization;          /* 48 */
    float surface_tension_coefficient;     /* 52 */
    float surface_tension_term;            /* 56 */
    float surface_tension_normalization;   /* 60 */
    float viscosity_constant;              /* 64 */
    float gravity;                         /* 68 */
    float delta_t;                         /* 72 */
    float density_variation_scaling_factor;/* 76 */
} Simulation_Params;
#pragma pack(pop)

#ifdef __cplusplus
static_assert(sizeof(Simulation_Params) == 80, "Simulation_Params must be 80 bytes");
static_assert(__builtin_offsetof(Simulation_Params, bucket_count) == 28, "layout");
static_assert(__builtin_offsetof(Simulation_Params, density_variation_scaling_factor) == 76, "layout");
}
#endif

#endif /* PCISPH_PARAMS_H */
