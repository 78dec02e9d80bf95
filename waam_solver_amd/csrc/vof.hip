// vof.hip — volume-of-fluid phase field: donor-cell advection, re-initialisation, the
// contact-angle ghost BC, flag classification, phi/f_l consistency and the freeze/remelt
// state flips.
#include "vof_cell.h"

namespace waam {

/* advect_phi — reference kernels.py:2320-2361 */
__global__ void k_advect_phi(float *__restrict__ dst, const float *__restrict__ src,
                             const float *__restrict__ ux, const float *__restrict__ uy,
                             const float *__restrict__ uz, const uint8_t *__restrict__ flags,
                             Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    size_t c = d.idx(i, j, k);
    dst[c] = advected_phi(src, ux, uy, uz, flags, i, j, k, d);
}

/* reinitialize_phi — reference kernels.py:2364-2391 */
__global__ void k_reinit(float *__restrict__ phi, const uint8_t *__restrict__ flags, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    phi[c] = reinit_phi(phi[c], flags[c]);
}

/* apply_contact_angle_phi_bc — reference kernels.py:933-991, Jacobi form */
__global__ void k_contact_bc(const float *__restrict__ pin, float *__restrict__ pout,
                             const uint8_t *__restrict__ flags, float cos_t, float sin_t, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    size_t c = d.idx(i, j, k);
    pout[c] = contact_bc_phi(pin, flags, i, j, k, cos_t, sin_t, d);
}

/* update_flags_from_phi — reference kernels.py:2394-2419 */
__global__ void k_update_flags(const float *__restrict__ phi, const float *__restrict__ f_l,
                               uint8_t *__restrict__ flags, int nz_solid, Dims d)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.n()) return;
    int i, j, k;
    d.ijk(c, i, j, k);
    flags[c] = classify_cell(k, nz_solid, phi[c], f_l[c]);
}

/* sync_phi_liquid_fraction — reference kernels.py:3484-3504 */
__global__ void k_sync_fl(const float *__restrict__ phi, float *__restrict__ f_l,
                          const uint8_t *__restrict__ flags, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    uint8_t f = flags[c];
    if (f == GAS || phi[c] < 0.05f) f_l[c] = 0.0f;
    else if (f == SOLID) f_l[c] = 0.0f;
}

/* solidify_cooled_metal — reference kernels.py:2422-2450 */
__global__ void k_solidify_cooled(const float *__restrict__ T, float *__restrict__ f_l,
                                  float *__restrict__ phi, uint8_t *__restrict__ flags,
                                  float *__restrict__ ux, float *__restrict__ uy,
                                  float *__restrict__ uz, float T_s, int zero_velocity, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N || flags[c] == GAS) return;
    if (phi[c] < 0.85f) return;
    if (f_l[c] < 0.02f && T[c] < T_s) {
        flags[c] = SOLID;
        phi[c] = 1.0f;
        f_l[c] = 0.0f;
        if (zero_velocity == 1) {
            ux[c] = 0.0f;
            uy[c] = 0.0f;
            uz[c] = 0.0f;
        }
    }
}

/* remelt_hot_solid_scalar — reference kernels.py:2489-2520 */
__global__ void k_remelt(const float *__restrict__ T, float *__restrict__ H,
                         float *__restrict__ f_l, float *__restrict__ phi,
                         uint8_t *__restrict__ flags, float cp_rho, float T_s, float T_l,
                         float L_rho, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N || flags[c] != SOLID) return;
    float T_c = T[c];
    if (T_c <= T_s + 5.0f) return;
    float H_sol = cp_rho * T_s;
    float fl;
    if (T_c >= T_l) {
        fl = 1.0f;
        H[c] = cp_rho * T_l + L_rho + cp_rho * (T_c - T_l);
    } else {
        fl = (T_c - T_s) / (T_l - T_s + 1e-6f);
        H[c] = H_sol + fl * L_rho;
    }
    f_l[c] = fl;
    phi[c] = 1.0f;
    flags[c] = FLUID;
}

/* solidify_trailing_pool_scalar — reference kernels.py:2573-2617 */
__global__ void k_trailing(float *__restrict__ T, float *__restrict__ H, float *__restrict__ f_l,
                           float *__restrict__ phi, uint8_t *__restrict__ flags,
                           float *__restrict__ ux, float *__restrict__ uy,
                           float *__restrict__ uz, float cp_rho, float arc_i, float arc_j,
                           float arc_k, float dx_, float dy_, float dz_, float lookback,
                           float T_freeze, float T_s, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    size_t c = d.idx(i, j, k);
    if (flags[c] == GAS) return;
    if (phi[c] < 0.85f || f_l[c] < 0.05f) return;
    float di = (float)i - arc_i, dj = (float)j - arc_j, dk = (float)k - arc_k;
    float proj = di * dx_ + dj * dy_ + dk * dz_;
    if (proj > -lookback) return;
    if (T[c] >= T_freeze) return;
    H[c] = cp_rho * (T_s - 1.0f);
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

int waam_advect_phi(float *phi_dst, const float *phi_src, const float *ux, const float *uy,
                    const float *uz, const uint8_t *flags, int nx, int ny, int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_advect_phi", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_advect_phi", {phi_dst, phi_src, ux, uy, uz, flags}));
    if (phi_dst == phi_src)
        return set_error(WAAM_ERR_INVALID, "waam_advect_phi: phi_dst aliases phi_src");
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_advect_phi, grid3(d), block3(), 0, as_stream(stream), phi_dst, phi_src,
                       ux, uy, uz, flags, d);
    return check_launch("waam_advect_phi");
}

int waam_reinitialize_phi(float *phi, const uint8_t *flags, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_reinitialize_phi: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_reinitialize_phi", {phi, flags}));
    hipLaunchKernelGGL(k_reinit, dim3(grid_for((size_t)n)), dim3(kBlock), 0, as_stream(stream),
                       phi, flags, (size_t)n);
    return check_launch("waam_reinitialize_phi");
}

int waam_apply_contact_angle_phi_bc(const float *phi_in, float *phi_out, const uint8_t *flags,
                                    float cos_t, float sin_t, int nx, int ny, int nz, int layout,
                                    void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_apply_contact_angle_phi_bc", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_apply_contact_angle_phi_bc", {phi_in, phi_out, flags}));
    if (phi_in == phi_out)
        return set_error(WAAM_ERR_INVALID, "waam_apply_contact_angle_phi_bc: phi_in aliases phi_out");
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_contact_bc, grid3(d), block3(), 0, as_stream(stream), phi_in, phi_out,
                       flags, cos_t, sin_t, d);
    return check_launch("waam_apply_contact_angle_phi_bc");
}

int waam_update_flags_from_phi(const float *phi, const float *f_l, uint8_t *flags, int nz_solid,
                               int nx, int ny, int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_update_flags_from_phi", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_update_flags_from_phi", {phi, f_l, flags}));
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_update_flags, dim3(grid_for(d.n())), dim3(kBlock), 0, as_stream(stream),
                       phi, f_l, flags, nz_solid, d);
    return check_launch("waam_update_flags_from_phi");
}

int waam_sync_phi_liquid_fraction(const float *phi, float *f_l, const uint8_t *flags, int64_t n,
                                  void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_sync_phi_liquid_fraction: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_sync_phi_liquid_fraction", {phi, f_l, flags}));
    hipLaunchKernelGGL(k_sync_fl, dim3(grid_for((size_t)n)), dim3(kBlock), 0, as_stream(stream),
                       phi, f_l, flags, (size_t)n);
    return check_launch("waam_sync_phi_liquid_fraction");
}

int waam_solidify_cooled_metal(const float *T, float *f_l, float *phi, uint8_t *flags, float *ux,
                               float *uy, float *uz, float T_solidus, int zero_velocity,
                               int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_solidify_cooled_metal: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_solidify_cooled_metal", {T, f_l, phi, flags, ux, uy, uz}));
    hipLaunchKernelGGL(k_solidify_cooled, dim3(grid_for((size_t)n)), dim3(kBlock), 0,
                       as_stream(stream), T, f_l, phi, flags, ux, uy, uz, T_solidus,
                       zero_velocity, (size_t)n);
    return check_launch("waam_solidify_cooled_metal");
}

int waam_remelt_hot_solid_scalar(const float *T, float *H, float *f_l, float *phi,
                                 uint8_t *flags, float cp_rho, float T_solidus,
                                 float T_liquidus, float L_rho, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_remelt_hot_solid_scalar: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_remelt_hot_solid_scalar", {T, H, f_l, phi, flags}));
    hipLaunchKernelGGL(k_remelt, dim3(grid_for((size_t)n)), dim3(kBlock), 0, as_stream(stream), T,
                       H, f_l, phi, flags, cp_rho, T_solidus, T_liquidus, L_rho, (size_t)n);
    return check_launch("waam_remelt_hot_solid_scalar");
}

int waam_solidify_trailing_pool_scalar(float *T, float *H, float *f_l, float *phi,
                                       uint8_t *flags, float *ux, float *uy, float *uz,
                                       float cp_rho, float arc_i, float arc_j, float arc_k,
                                       float dir_x, float dir_y, float dir_z,
                                       float lookback_cells, float T_freeze, float T_solidus,
                                       int nx, int ny, int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_solidify_trailing_pool_scalar", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_solidify_trailing_pool_scalar", {T, H, f_l, phi, flags, ux, uy, uz}));
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_trailing, grid3(d), block3(), 0, as_stream(stream), T, H, f_l, phi, flags,
                       ux, uy, uz, cp_rho, arc_i, arc_j, arc_k, dir_x, dir_y, dir_z,
                       lookback_cells, T_freeze, T_solidus, d);
    return check_launch("waam_solidify_trailing_pool_scalar");
}

}  // extern "C"
