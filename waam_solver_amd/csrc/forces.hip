// forces.hip — standalone body-force kernels (each adds its term with +=, as the reference).
#include "forces_cell.h"

namespace waam {

#define CELL3D                    \
    int i, j, k;                  \
    if (!d.cell(i, j, k)) return; \
    size_t c = d.idx(i, j, k);

__global__ void k_csf(const float *__restrict__ phi, const uint8_t *__restrict__ flags,
                      float *__restrict__ Fx, float *__restrict__ Fy, float *__restrict__ Fz,
                      float gamma_lu, int wetting, float cos_t, float sin_t, Dims d)
{
    CELL3D
    uint8_t f = flags[c];
    if (f == SOLID || f == GAS) return;
    ForceStencil fs;
    load_force_stencil(phi, phi, flags, i, j, k, d, fs);
    float fx, fy, fz;
    if (!csf_force(fs, i, j, k, d, gamma_lu, wetting, cos_t, sin_t, fx, fy, fz)) return;
    Fx[c] += fx;
    Fy[c] += fy;
    Fz[c] += fz;
}

/* compute_curvature_field — reference kernels.py:3353-3376 */
__global__ void k_curvature(const float *__restrict__ phi, const uint8_t *__restrict__ flags,
                            float *__restrict__ kappa_out, Dims d)
{
    CELL3D
    kappa_out[c] = 0.0f;
    if (flags[c] == GAS) return;
    if (i < 1 || j < 1 || k < 1 || i > d.nx - 2 || j > d.ny - 2 || k > d.nz - 2) return;
    PhiStencil st;
    load_phi_stencil(phi, i, j, k, d, st);
    float kappa, g;
    brackbill_kappa(st, 1e-6f, kappa, g);
    if (g >= 1e-6f) kappa_out[c] = kappa;
}

__global__ void k_marangoni(const float *__restrict__ T, const float *__restrict__ phi,
                            const float *__restrict__ f_l, float *__restrict__ Fx,
                            float *__restrict__ Fy, float *__restrict__ Fz,
                            const uint8_t *__restrict__ flags, float dgamma, Dims d)
{
    CELL3D
    uint8_t f = flags[c];
    if (f == SOLID || f == GAS) return;
    ForceStencil fs;
    load_force_stencil(phi, T, flags, i, j, k, d, fs);
    float fx, fy, fz;
    if (!marangoni_force(fs, f_l[c], dgamma, fx, fy, fz)) return;
    Fx[c] += fx;
    Fy[c] += fy;
    Fz[c] += fz;
}

/* add_buoyancy — reference kernels.py:2748-2753 */
__global__ void k_buoyancy(const float *__restrict__ T, float *__restrict__ Fz,
                           const float *__restrict__ f_l, const uint8_t *__restrict__ flags,
                           float g_lu, float beta, float T_ref, float rho_ref, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    uint8_t f = flags[c];
    if (f == SOLID || f == GAS) return;
    float fl = f_l[c];
    if (fl > 0.0f) Fz[c] += rho_ref * g_lu * beta * (T[c] - T_ref) * fl;
}

/* add_hydrostatic_gravity — reference kernels.py:2767-2772 */
__global__ void k_hydrostatic(float *__restrict__ Fz, const float *__restrict__ f_l,
                              const uint8_t *__restrict__ flags, float rho_ref, float g_lu,
                              size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    uint8_t f = flags[c];
    if (f == SOLID || f == GAS) return;
    float fl = f_l[c];
    if (fl > 0.0f) Fz[c] += -rho_ref * g_lu * fl;
}

__global__ void k_arc_pressure(float *__restrict__ Fz, const uint8_t *__restrict__ flags,
                               const float *__restrict__ phi, ArcPressure ap, Dims d)
{
    CELL3D
    uint8_t f = flags[c];
    if (f == SOLID || f == GAS) return;
    float fz;
    float pzp = phi[d.idx(i, j, imin(k + 1, d.nz - 1))], pzm = phi[d.idx(i, j, imax(k - 1, 0))];
    if (arc_pressure_force(pzp, pzm, i, j, k, ap, fz)) Fz[c] += fz;
}

/* apply_vapor_recoil / apply_vapor_recoil_clausius_clapeyron — reference kernels.py:995-1030,
 * 1052-1090 */
__global__ void k_vapor_recoil(float *__restrict__ Fz, const float *__restrict__ T,
                               const float *__restrict__ phi, const uint8_t *__restrict__ flags,
                               Recoil r, Dims d)
{
    CELL3D
    uint8_t f = flags[c];
    if (f == SOLID || f == GAS) return;
    float pzp = phi[d.idx(i, j, imin(k + 1, d.nz - 1))], pzm = phi[d.idx(i, j, imax(k - 1, 0))];
    float fz;
    if (recoil_force(pzp, pzm, T[c], i, j, r, fz)) Fz[c] += fz;
}

/* apply_gas_shear_stress — reference kernels.py:1093-1139 */
__global__ void k_gas_shear(float *__restrict__ Fx, float *__restrict__ Fy,
                            const float *__restrict__ phi, const uint8_t *__restrict__ flags,
                            GasShear g, Dims d)
{
    CELL3D
    uint8_t f = flags[c];
    if (f == SOLID || f == GAS) return;
    PhiStencil ps;
    size_t o[6];
    face_idx(i, j, k, d, o);
#pragma unroll
    for (int q = 0; q < 6; ++q) ps.c[q] = phi[o[q]];
    float fx, fy;
    if (gas_shear_force(ps, i, j, g, fx, fy)) {
        Fx[c] += fx;
        Fy[c] += fy;
    }
}

/* clamp_body_force_magnitude — reference kernels.py:3298-3309 */
__global__ void k_clamp_force(float *__restrict__ Fx, float *__restrict__ Fy,
                              float *__restrict__ Fz, const uint8_t *__restrict__ flags,
                              float F_max, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    uint8_t f = flags[c];
    if (f == SOLID || f == GAS) return;
    float fx = Fx[c], fy = Fy[c], fz = Fz[c];
    float mag = sqrtf(fx * fx + fy * fy + fz * fz);
    if (mag > F_max) {
        float s = F_max / (mag + 1e-12f);
        Fx[c] = fx * s;
        Fy[c] = fy * s;
        Fz[c] = fz * s;
    }
}

/* apply_droplet_impact_pressure — reference kernels.py:1159-1173 */
__global__ void k_droplet_pressure(float *__restrict__ Fz, const uint8_t *__restrict__ flags,
                                   const float *__restrict__ phi, const float *__restrict__ f_l,
                                   float arc_i, float arc_j, float arc_k, float dr, float inv2s2,
                                   float F_peak_lu, Dims d, const float *__restrict__ arc_k_dev)
{
    CELL3D
    if (arc_k_dev) arc_k = *arc_k_dev;  // resolved on the device
    if (flags[c] == GAS) return;
    if (f_l[c] < 0.25f && phi[c] < 0.25f) return;
    float di = (float)i - arc_i, dj = (float)j - arc_j, dk = (float)k - arc_k;
    float r2 = di * di + dj * dj + dk * dk;
    if (r2 > dr * dr * 4.0f) return;
    Fz[c] += -F_peak_lu * waam_expf(-r2 * inv2s2);
}

}  // namespace waam

using namespace waam;

extern "C" {

int waam_compute_csf_tension(const float *phi, const uint8_t *flags, float *Fx, float *Fy,
                             float *Fz, float gamma_lu, int enable_wetting, float cos_t,
                             float sin_t, int nx, int ny, int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_compute_csf_tension", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_compute_csf_tension", {phi, flags, Fx, Fy, Fz}));
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_csf, grid3(d), block3(), 0, as_stream(stream), phi, flags, Fx, Fy, Fz,
                       gamma_lu, enable_wetting, cos_t, sin_t, d);
    return check_launch("waam_compute_csf_tension");
}

int waam_compute_curvature_field(const float *phi, const uint8_t *flags, float *kappa_out,
                                 int nx, int ny, int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_compute_curvature_field", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_compute_curvature_field", {phi, flags, kappa_out}));
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_curvature, grid3(d), block3(), 0, as_stream(stream), phi, flags,
                       kappa_out, d);
    return check_launch("waam_compute_curvature_field");
}

int waam_compute_marangoni_force(const float *T, const float *phi, const float *f_l, float *Fx,
                                 float *Fy, float *Fz, const uint8_t *flags, float dgamma_dT,
                                 int nx, int ny, int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_compute_marangoni_force", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_compute_marangoni_force", {T, phi, f_l, Fx, Fy, Fz, flags}));
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_marangoni, grid3(d), block3(), 0, as_stream(stream), T, phi, f_l, Fx, Fy,
                       Fz, flags, dgamma_dT, d);
    return check_launch("waam_compute_marangoni_force");
}

int waam_add_buoyancy(const float *T, float *Fz, const float *f_l, const uint8_t *flags,
                      float g_lu, float beta, float T_ref, float rho_ref, int64_t n,
                      void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_add_buoyancy: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_add_buoyancy", {T, Fz, f_l, flags}));
    hipLaunchKernelGGL(k_buoyancy, dim3(grid_for((size_t)n)), dim3(kBlock), 0, as_stream(stream),
                       T, Fz, f_l, flags, g_lu, beta, T_ref, rho_ref, (size_t)n);
    return check_launch("waam_add_buoyancy");
}

int waam_add_hydrostatic_gravity(float *Fz, const float *f_l, const uint8_t *flags,
                                 float rho_ref, float g_lu, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_add_hydrostatic_gravity: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_add_hydrostatic_gravity", {Fz, f_l, flags}));
    hipLaunchKernelGGL(k_hydrostatic, dim3(grid_for((size_t)n)), dim3(kBlock), 0,
                       as_stream(stream), Fz, f_l, flags, rho_ref, g_lu, (size_t)n);
    return check_launch("waam_add_hydrostatic_gravity");
}

int waam_apply_arc_pressure(float *Fz, const uint8_t *flags, const float *phi, float arc_i,
                            float arc_j, float arc_k, float sigma, float pressure_pa, float dt,
                            float dx, float rho_ref, int nx, int ny, int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_apply_arc_pressure", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_apply_arc_pressure", {Fz, flags, phi}));
    Dims d{nx, ny, nz, layout};
    ArcPressure ap;
    ap.arc_i = arc_i;
    ap.arc_j = arc_j;
    ap.arc_k = arc_k;
    ap.inv2s2 = 1.0f / (2.0f * sigma * sigma + 1e-6f);
    float F_peak_phys = pressure_pa / (rho_ref * dx);
    ap.F_peak_lu = F_peak_phys * (dt * dt) / dx;
    hipLaunchKernelGGL(k_arc_pressure, grid3(d), block3(), 0, as_stream(stream), Fz, flags, phi, ap,
                       d);
    return check_launch("waam_apply_arc_pressure");
}

static Recoil recoil_params(float arc_i, float arc_j, float sigma)
{
    Recoil r{};
    r.arc_i = arc_i;
    r.arc_j = arc_j;
    r.inv2s2 = 1.0f / (2.0f * sigma * sigma + 1e-6f);
    return r;
}

int waam_apply_vapor_recoil(float *Fz, const float *T, const float *phi, const uint8_t *flags,
                            float arc_i, float arc_j, float arc_k, float sigma, float recoil_pa,
                            float T_ref, float dt, float dx, float rho_ref, int nx, int ny,
                            int nz, int layout, void *stream)
{
    (void)arc_k;
    WAAM_RETURN_IF(check_dims("waam_apply_vapor_recoil", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_apply_vapor_recoil", {Fz, T, phi, flags}));
    Dims d{nx, ny, nz, layout};
    Recoil r = recoil_params(arc_i, arc_j, sigma);
    r.clausius = 0;
    r.F_peak_lu = recoil_pa / (rho_ref * dx) * dt * dt / dx;
    r.T_ref = T_ref;
    hipLaunchKernelGGL(k_vapor_recoil, grid3(d), block3(), 0, as_stream(stream), Fz, T, phi,
                       flags, r, d);
    return check_launch("waam_apply_vapor_recoil");
}

int waam_apply_vapor_recoil_clausius_clapeyron(float *Fz, const float *T, const float *phi,
                                               const uint8_t *flags, float arc_i, float arc_j,
                                               float arc_k, float sigma, float P_ref_Pa,
                                               float T_boil_K, float L_vapor_J_kg,
                                               float R_spec_J_kgK, float C_acc, float dt, float dx,
                                               float rho_ref, int nx, int ny, int nz, int layout,
                                               void *stream)
{
    (void)arc_k;
    WAAM_RETURN_IF(check_dims("waam_apply_vapor_recoil_clausius_clapeyron", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_apply_vapor_recoil_clausius_clapeyron", {Fz, T, phi, flags}));
    Dims d{nx, ny, nz, layout};
    Recoil r = recoil_params(arc_i, arc_j, sigma);
    r.clausius = 1;
    r.P_ref = P_ref_Pa;
    r.T_boil = T_boil_K;
    r.L_vap = L_vapor_J_kg;
    r.R_spec = R_spec_J_kgK;
    r.C_acc = C_acc;
    r.dt = dt;
    r.dx = dx;
    r.rho_ref = rho_ref;
    hipLaunchKernelGGL(k_vapor_recoil, grid3(d), block3(), 0, as_stream(stream), Fz, T, phi,
                       flags, r, d);
    return check_launch("waam_apply_vapor_recoil_clausius_clapeyron");
}

int waam_apply_gas_shear_stress(float *Fx, float *Fy, const float *phi, const uint8_t *flags,
                                float arc_i, float arc_j, float arc_k, float sigma,
                                float tau_peak_pa, float dt, float dx, float rho_ref, int nx,
                                int ny, int nz, int layout, void *stream)
{
    (void)arc_k;
    WAAM_RETURN_IF(check_dims("waam_apply_gas_shear_stress", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_apply_gas_shear_stress", {Fx, Fy, phi, flags}));
    Dims d{nx, ny, nz, layout};
    GasShear g;
    g.arc_i = arc_i;
    g.arc_j = arc_j;
    g.inv2s2 = 1.0f / (2.0f * sigma * sigma + 1e-6f);
    g.tau_peak = tau_peak_pa;
    g.F_scale = 1.0f / (rho_ref * dx) * dt * dt / dx;
    hipLaunchKernelGGL(k_gas_shear, grid3(d), block3(), 0, as_stream(stream), Fx, Fy, phi, flags,
                       g, d);
    return check_launch("waam_apply_gas_shear_stress");
}

int waam_clamp_body_force_magnitude(float *Fx, float *Fy, float *Fz, const uint8_t *flags,
                                    float F_max_lu, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_clamp_body_force_magnitude: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_clamp_body_force_magnitude", {Fx, Fy, Fz, flags}));
    hipLaunchKernelGGL(k_clamp_force, dim3(grid_for((size_t)n)), dim3(kBlock), 0,
                       as_stream(stream), Fx, Fy, Fz, flags, F_max_lu, (size_t)n);
    return check_launch("waam_clamp_body_force_magnitude");
}

int waam_apply_droplet_impact_pressure(float *Fz, const uint8_t *flags, const float *phi,
                                       const float *f_l, float arc_i, float arc_j, float arc_k,
                                       float drop_radius, float impact_pa, float dt, float dx,
                                       float rho_ref, int nx, int ny, int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_apply_droplet_impact_pressure", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_apply_droplet_impact_pressure", {Fz, flags, phi, f_l}));
    Dims d{nx, ny, nz, layout};
    float inv2s2 = 1.0f / (2.0f * drop_radius * drop_radius + 1e-6f);
    float F_peak_phys = impact_pa / (rho_ref * dx);
    float F_peak_lu = F_peak_phys * dt * dt / dx;
    hipLaunchKernelGGL(k_droplet_pressure, grid3(d), block3(), 0, as_stream(stream), Fz, flags,
                       phi, f_l, arc_i, arc_j, arc_k, drop_radius, inv2s2, F_peak_lu, d,
                       (const float *)nullptr);
    return check_launch("waam_apply_droplet_impact_pressure");
}

}  // extern "C"

namespace waam {

int launch_droplet_pressure_dev(float *Fz, const uint8_t *flags, const float *phi,
                                const float *f_l, float arc_i, float arc_j,
                                const float *arc_k_dev, float drop_radius, float impact_pa,
                                float dt, float dx, float rho_ref, Dims d, hipStream_t st)
{
    float inv2s2 = 1.0f / (2.0f * drop_radius * drop_radius + 1e-6f);
    float F_peak_phys = impact_pa / (rho_ref * dx);
    float F_peak_lu = F_peak_phys * dt * dt / dx;
    hipLaunchKernelGGL(k_droplet_pressure, grid3(d), block3(), 0, st, Fz, flags, phi, f_l, arc_i,
                       arc_j, 0.0f, drop_radius, inv2s2, F_peak_lu, d, arc_k_dev);
    return check_launch("waam_step_fused_droplet (impact pressure)");
}

}  // namespace waam
