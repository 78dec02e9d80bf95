// droplet_launch.h — launches of the droplet block with arc_k resolved on the device (read from
// a device float instead of passed by the host), used by the fused droplet step.
#pragma once

#include "waam_common.h"

namespace waam {

int launch_deposit_dev(float *f_src, uint8_t *flags, float *f_l, float *phi, float *H, float *T,
                       float *rho, float *vol_acc, float *real_acc, void *scratch, float *result,
                       float arc_i, float arc_j, const float *arc_k_dev, double footprint_r,
                       float droplet_radius, double drop_vol, float T_drop, float cp_rho,
                       float L_rho, float rho0, float cell_vol, Dims d, hipStream_t st);
int launch_tracers_dev(float *pos, int32_t *active, int32_t *head, int max_tracers, float ax,
                       float ay, const float *arc_k_dev, double dx, float sigma_m, int spawn,
                       uint32_t seed, hipStream_t st);
int launch_momentum_impact_dev(float *f_src, const float *rho, float *ux, float *uy, float *uz,
                               const uint8_t *flags, const float *f_l, float arc_i, float arc_j,
                               const float *arc_k_dev, float dr, float vx, float vy, float vz,
                               Dims d, hipStream_t st);
int launch_droplet_pressure_dev(float *Fz, const uint8_t *flags, const float *phi,
                                const float *f_l, float arc_i, float arc_j,
                                const float *arc_k_dev, float drop_radius, float impact_pa,
                                float dt, float dx, float rho_ref, Dims d, hipStream_t st);

}  // namespace waam
