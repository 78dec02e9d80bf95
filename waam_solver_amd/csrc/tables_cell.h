// tables_cell.h — per-cell material properties from piecewise-linear tables (reference
// kernels.py:1685-1772), shared by the refresh kernel and the fused step.
#pragma once

#include "waam_common.h"

namespace waam {

/* lookup_table_1d — reference kernels.py:1685-1707: clamp outside the knots; inside, every
 * bracketing interval overwrites the result (the last one wins, which matters on a knot). */
__device__ __forceinline__ float lookup_table_1d(float T, const float *t_arr, const float *v_arr,
                                                 int n, float fallback)
{
    float result = fallback;
    if (n > 0) {
        if (T <= t_arr[0]) {
            result = v_arr[0];
        } else if (T >= t_arr[n - 1]) {
            result = v_arr[n - 1];
        } else {
#pragma unroll
            for (int idx = 0; idx < WAAM_MAX_KNOTS - 1; ++idx) {
                if (idx < n - 1) {
                    float t0 = t_arr[idx], t1 = t_arr[idx + 1];
                    if (T >= t0 && T <= t1) {
                        float w = (T - t0) / (t1 - t0 + 1e-8f);
                        result = v_arr[idx] + w * (v_arr[idx + 1] - v_arr[idx]);
                    }
                }
            }
        }
    }
    return result;
}

/* refresh_thermal_properties for one cell (reference kernels.py:1710-1772): the constants on
 * SOLID and GAS cells, else cp*rho, the lattice diffusivity, the Marangoni coefficient and the
 * relaxation time from the tables at the cell's temperature (use_tables 0: the constants, the
 * Marangoni one scaled) */
struct CellProps {
    float cp_rho, alpha_lu, dgamma_lu, tau;
};

__device__ __forceinline__ CellProps cell_props(const waam_material_tables &tb,
                                                const waam_property_refs &rf, int use_tables,
                                                uint8_t flag, float T_c)
{
    CellProps pr;
    if (flag == SOLID || flag == GAS) {
        pr.cp_rho = rf.cp_rho_ref;
        pr.alpha_lu = rf.alpha_lu_ref;
        pr.dgamma_lu = rf.dgamma_lu_ref;
        pr.tau = rf.tau_ref;
    } else if (use_tables == 1) {
        float cp = lookup_table_1d(T_c, tb.cp_T, tb.cp_V, tb.n_cp, tb.cp_fallback);
        float k_val = lookup_table_1d(T_c, tb.k_T, tb.k_V, tb.n_k, tb.k_fallback);
        float mu_val = lookup_table_1d(T_c, tb.mu_T, tb.mu_V, tb.n_mu, tb.mu_fallback);
        float dgamma = lookup_table_1d(T_c, tb.dgamma_T, tb.dgamma_V, tb.n_dgamma,
                                       tb.dgamma_fallback);
        float alpha_phys = k_val / (rf.rho * cp + 1e-9f);
        float nu_phys = mu_val / (rf.rho + 1e-9f);
        float nu_lu = nu_phys * rf.dt / (rf.dx * rf.dx);
        pr.cp_rho = rf.rho * cp;
        pr.alpha_lu = alpha_phys * rf.dt / (rf.dx * rf.dx);
        pr.dgamma_lu = dgamma * rf.force_scale * rf.marangoni_scale;
        pr.tau = 3.0f * nu_lu + 0.5f;
    } else {
        pr.cp_rho = rf.cp_rho_ref;
        pr.alpha_lu = rf.alpha_lu_ref;
        pr.dgamma_lu = rf.dgamma_lu_ref * rf.marangoni_scale;
        pr.tau = rf.tau_ref;
    }
    return pr;
}

}  // namespace waam
