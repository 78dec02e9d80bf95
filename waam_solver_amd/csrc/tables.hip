// tables.hip — the temperature-dependent material path: per-cell properties from piecewise-
// linear tables and the kernels that read them (reference kernels.py:1685-2010, 2220-2269,
// 2454-2486, 2524-2570). Each kernel equals its constant-property counterpart with the
// constant replaced by the cell's value from the property fields.
#include "forces_cell.h"
#include "tables_cell.h"
#include "thermal_cell.h"

namespace waam {

/* refresh_thermal_properties — reference kernels.py:1710-1772 */
__global__ void k_refresh_properties(const float *__restrict__ T, float *__restrict__ cp_rho_f,
                                     float *__restrict__ alpha_f, float *__restrict__ dgamma_f,
                                     float *__restrict__ tau_f, waam_material_tables tb,
                                     waam_property_refs rf, int use_tables,
                                     const uint8_t *__restrict__ flags, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const CellProps pr = cell_props(tb, rf, use_tables, flags[c], T[c]);
    cp_rho_f[c] = pr.cp_rho;
    alpha_f[c] = pr.alpha_lu;
    dgamma_f[c] = pr.dgamma_lu;
    tau_f[c] = pr.tau;
}

/* advect_diffuse_temperature_variable — reference kernels.py:1775-1836 (Jacobi, as the
 * constant-property kernel) */
__global__ void k_advect_diffuse_variable(const float *__restrict__ H_in,
                                          float *__restrict__ H_out, const float *__restrict__ T,
                                          const float *__restrict__ ux,
                                          const float *__restrict__ uy,
                                          const float *__restrict__ uz,
                                          const uint8_t *__restrict__ flags,
                                          const float *__restrict__ alpha_f,
                                          const float *__restrict__ cp_rho_f, float dt, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    size_t c = d.idx(i, j, k);
    uint8_t flag = flags[c];
    if (flag == GAS) {
        H_out[c] = H_in[c];
        return;
    }
    H_out[c] = H_in[c] + thermal_dH(H_in, T, ux, uy, uz, flags, i, j, k, flag, alpha_f[c], dt,
                                    cp_rho_f[c], d);
}

/* update_phase_variable_cp — reference kernels.py:1839-1863 */
__global__ void k_phase_variable(const float *__restrict__ H, float *__restrict__ T,
                                 float *__restrict__ f_l, const float *__restrict__ cp_rho_f,
                                 float L_rho, float T_s, float T_l, float H_sol, float H_liq,
                                 size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float h = H[c], cp_rho = cp_rho_f[c];
    if (h <= H_sol) {
        f_l[c] = 0.0f;
        T[c] = h / (cp_rho + 1e-9f);
    } else if (h >= H_liq) {
        f_l[c] = 1.0f;
        T[c] = T_l + (h - H_liq) / (cp_rho + 1e-9f);
    } else {
        float fl = (h - H_sol) / L_rho;
        f_l[c] = fl;
        T[c] = T_s + fl * (T_l - T_s);
    }
}

/* clamp_enthalpy_floor — reference kernels.py:1937-1950 */
__global__ void k_floor_variable(float *__restrict__ H, const float *__restrict__ cp_rho_f,
                                 const uint8_t *__restrict__ flags, float T_floor, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N || flags[c] == GAS) return;
    float h_min = cp_rho_f[c] * T_floor;
    if (H[c] < h_min) H[c] = h_min;
}

/* clamp_enthalpy_ceiling_variable_cp — reference kernels.py:1987-2010 */
__global__ void k_ceiling_variable(float *__restrict__ H, const uint8_t *__restrict__ flags,
                                   const float *__restrict__ cp_rho_f, float H_liq, float T_l,
                                   float T_cap, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N || flags[c] == GAS) return;
    float cp_r = cp_rho_f[c];
    H[c] = ceiling_clamp(H[c], H_liq, H_liq + cp_r * (T_cap - T_l), cp_r * T_cap);
}

/* compute_marangoni_force_variable — reference kernels.py:2220-2269 */
__global__ void k_marangoni_variable(const float *__restrict__ T, const float *__restrict__ phi,
                                     const float *__restrict__ f_l, float *__restrict__ Fx,
                                     float *__restrict__ Fy, float *__restrict__ Fz,
                                     const uint8_t *__restrict__ flags,
                                     const float *__restrict__ dgamma_f, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    size_t c = d.idx(i, j, k);
    uint8_t f = flags[c];
    if (f == SOLID || f == GAS) return;
    ForceStencil fs;
    load_force_stencil(phi, T, flags, i, j, k, d, fs);
    float fx, fy, fz;
    if (!marangoni_force(fs, f_l[c], dgamma_f[c], fx, fy, fz)) return;
    Fx[c] += fx;
    Fy[c] += fy;
    Fz[c] += fz;
}

/* remelt_hot_solid — reference kernels.py:2454-2486 */
__global__ void k_remelt_variable(const float *__restrict__ T, float *__restrict__ H,
                                  float *__restrict__ f_l, float *__restrict__ phi,
                                  uint8_t *__restrict__ flags, const float *__restrict__ cp_rho_f,
                                  float T_s, float T_l, float L_rho, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N || flags[c] != SOLID) return;
    float T_c = T[c];
    if (T_c <= T_s + 5.0f) return;
    float cp_r = cp_rho_f[c];
    float H_sol = cp_r * T_s, fl;
    if (T_c >= T_l) {
        fl = 1.0f;
        H[c] = cp_r * T_l + L_rho + cp_r * (T_c - T_l);
    } else {
        fl = (T_c - T_s) / (T_l - T_s + 1e-6f);
        H[c] = H_sol + fl * L_rho;
    }
    f_l[c] = fl;
    phi[c] = 1.0f;
    flags[c] = FLUID;
}

/* solidify_trailing_pool — reference kernels.py:2524-2570 */
__global__ void k_trailing_variable(float *__restrict__ T, float *__restrict__ H,
                                    float *__restrict__ f_l, float *__restrict__ phi,
                                    uint8_t *__restrict__ flags, float *__restrict__ ux,
                                    float *__restrict__ uy, float *__restrict__ uz,
                                    const float *__restrict__ cp_rho_f, float arc_i, float arc_j,
                                    float arc_k, float dir_x, float dir_y, float dir_z,
                                    float lookback, float T_freeze, float T_s, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    size_t c = d.idx(i, j, k);
    if (flags[c] == GAS) return;
    if (phi[c] < 0.85f || f_l[c] < 0.05f) return;
    float di = (float)i - arc_i, dj = (float)j - arc_j, dk = (float)k - arc_k;
    float proj = di * dir_x + dj * dir_y + dk * dir_z;
    if (proj > -lookback) return;
    if (T[c] >= T_freeze) return;
    float cp_r = cp_rho_f[c];
    H[c] = cp_r * (T_s - 1.0f);
    T[c] = T_s - 1.0f;
    f_l[c] = 0.0f;
    phi[c] = 1.0f;
    flags[c] = SOLID;
    ux[c] = 0.0f;
    uy[c] = 0.0f;
    uz[c] = 0.0f;
}

}  // namespace waam

using namespace waam;

extern "C" {

int waam_refresh_thermal_properties(const float *T, float *cp_rho_field, float *alpha_lu_field,
                                    float *dgamma_lu_field, float *tau_field,
                                    const waam_material_tables *tables,
                                    const waam_property_refs *refs, int use_tables,
                                    const uint8_t *flags, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_refresh_thermal_properties: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_refresh_thermal_properties",
                              {T, cp_rho_field, alpha_lu_field, dgamma_lu_field, tau_field, tables,
                               refs, flags}));
    const waam_material_tables &tb = *tables;
    if (tb.n_cp > WAAM_MAX_KNOTS || tb.n_k > WAAM_MAX_KNOTS || tb.n_mu > WAAM_MAX_KNOTS ||
        tb.n_dgamma > WAAM_MAX_KNOTS || tb.n_cp < 0 || tb.n_k < 0 || tb.n_mu < 0 || tb.n_dgamma < 0)
        return set_error(WAAM_ERR_INVALID, "waam_refresh_thermal_properties: knot count outside 0..%d",
                         WAAM_MAX_KNOTS);
    hipLaunchKernelGGL(k_refresh_properties, dim3(grid_for((size_t)n)), dim3(kBlock), 0,
                       as_stream(stream), T, cp_rho_field, alpha_lu_field, dgamma_lu_field,
                       tau_field, tb, *refs, use_tables, flags, (size_t)n);
    return check_launch("waam_refresh_thermal_properties");
}

int waam_advect_diffuse_temperature_variable(const float *H_in, float *H_out, const float *T,
                                             const float *ux, const float *uy, const float *uz,
                                             const uint8_t *flags, const float *alpha_lu_field,
                                             const float *cp_rho_field, float dt, int nx, int ny,
                                             int nz, int layout, void *stream)
{
    const char *name = "waam_advect_diffuse_temperature_variable";
    WAAM_RETURN_IF(check_dims(name, nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs(name, {H_in, H_out, T, ux, uy, uz, flags, alpha_lu_field, cp_rho_field}));
    if (H_in == H_out) return set_error(WAAM_ERR_INVALID, "%s: H_in aliases H_out", name);
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_advect_diffuse_variable, grid3(d), block3(), 0, as_stream(stream), H_in,
                       H_out, T, ux, uy, uz, flags, alpha_lu_field, cp_rho_field, dt, d);
    return check_launch(name);
}

int waam_update_phase_variable_cp(const float *H, float *T, float *f_l, const float *cp_rho_field,
                                  float L_rho, float T_solidus, float T_liquidus, float H_sol,
                                  float H_liq, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_update_phase_variable_cp: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_update_phase_variable_cp", {H, T, f_l, cp_rho_field}));
    hipLaunchKernelGGL(k_phase_variable, dim3(grid_for((size_t)n)), dim3(kBlock), 0,
                       as_stream(stream), H, T, f_l, cp_rho_field, L_rho, T_solidus, T_liquidus,
                       H_sol, H_liq, (size_t)n);
    return check_launch("waam_update_phase_variable_cp");
}

int waam_clamp_enthalpy_floor(float *H, const float *cp_rho_field, const uint8_t *flags,
                              float T_floor, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_clamp_enthalpy_floor: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_clamp_enthalpy_floor", {H, cp_rho_field, flags}));
    hipLaunchKernelGGL(k_floor_variable, dim3(grid_for((size_t)n)), dim3(kBlock), 0,
                       as_stream(stream), H, cp_rho_field, flags, T_floor, (size_t)n);
    return check_launch("waam_clamp_enthalpy_floor");
}

int waam_clamp_enthalpy_ceiling_variable_cp(float *H, const uint8_t *flags,
                                            const float *cp_rho_field, float H_liq,
                                            float T_liquidus, float T_vapor_cap, int64_t n,
                                            void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_clamp_enthalpy_ceiling_variable_cp: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_clamp_enthalpy_ceiling_variable_cp", {H, flags, cp_rho_field}));
    hipLaunchKernelGGL(k_ceiling_variable, dim3(grid_for((size_t)n)), dim3(kBlock), 0,
                       as_stream(stream), H, flags, cp_rho_field, H_liq, T_liquidus, T_vapor_cap,
                       (size_t)n);
    return check_launch("waam_clamp_enthalpy_ceiling_variable_cp");
}

int waam_compute_marangoni_force_variable(const float *T, const float *phi, const float *f_l,
                                          float *Fx, float *Fy, float *Fz, const uint8_t *flags,
                                          const float *dgamma_lu_field, int nx, int ny, int nz,
                                          int layout, void *stream)
{
    const char *name = "waam_compute_marangoni_force_variable";
    WAAM_RETURN_IF(check_dims(name, nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs(name, {T, phi, f_l, Fx, Fy, Fz, flags, dgamma_lu_field}));
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_marangoni_variable, grid3(d), block3(), 0, as_stream(stream), T, phi, f_l,
                       Fx, Fy, Fz, flags, dgamma_lu_field, d);
    return check_launch(name);
}

int waam_remelt_hot_solid(const float *T, float *H, float *f_l, float *phi, uint8_t *flags,
                          const float *cp_rho_field, float T_solidus, float T_liquidus,
                          float L_rho, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_remelt_hot_solid: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_remelt_hot_solid", {T, H, f_l, phi, flags, cp_rho_field}));
    hipLaunchKernelGGL(k_remelt_variable, dim3(grid_for((size_t)n)), dim3(kBlock), 0,
                       as_stream(stream), T, H, f_l, phi, flags, cp_rho_field, T_solidus,
                       T_liquidus, L_rho, (size_t)n);
    return check_launch("waam_remelt_hot_solid");
}

int waam_solidify_trailing_pool(float *T, float *H, float *f_l, float *phi, uint8_t *flags,
                                float *ux, float *uy, float *uz, const float *cp_rho_field,
                                float arc_i, float arc_j, float arc_k, float dir_x, float dir_y,
                                float dir_z, float lookback_cells, float T_freeze,
                                float T_solidus, int nx, int ny, int nz, int layout, void *stream)
{
    const char *name = "waam_solidify_trailing_pool";
    WAAM_RETURN_IF(check_dims(name, nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs(name, {T, H, f_l, phi, flags, ux, uy, uz, cp_rho_field}));
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_trailing_variable, grid3(d), block3(), 0, as_stream(stream), T, H, f_l,
                       phi, flags, ux, uy, uz, cp_rho_field, arc_i, arc_j, arc_k, dir_x, dir_y,
                       dir_z, lookback_cells, T_freeze, T_solidus, d);
    return check_launch(name);
}

}  // extern "C"
