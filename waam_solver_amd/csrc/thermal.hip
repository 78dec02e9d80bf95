// thermal.hip — arc heat input, enthalpy advection-diffusion, boundary losses, enthalpy clamps,
// the enthalpy -> (T, f_l) phase map and the heat-affected-zone diagnostics.
#include "arc_cell.h"
#include "thermal_cell.h"

namespace waam {

/* ---------------------------------------------------------------------------------------------
 * Arc heat (reference kernels.py:406-622). Pass 1 reduces the weight total; the reference uses
 * one float atomic per cell, here every thread keeps a compensated double sum, blocks combine
 * pairs with TwoSum and one block adds the 1024 block results in a fixed order, so the total is
 * the exact sum to within a couple of double ulps, independent of scheduling, rounded once to
 * fp32. Pass 2 recomputes each weight and adds energy_per_weight * w to H.
 * ------------------------------------------------------------------------------------------- */
__global__ void __launch_bounds__(kBlock) k_arc_partials(const uint8_t *__restrict__ flags,
                                                         const float *__restrict__ f_l,
                                                         ArcParams p, Dims d,
                                                         double *__restrict__ partials)
{
    __shared__ double sh_s[kBlock], sh_c[kBlock];
    size_t N = d.n();
    double s = 0.0, c = 0.0;
    for (size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x; cell < N;
         cell += (size_t)gridDim.x * blockDim.x) {
        if (flags[cell] == GAS) continue;
        int i, j, k;
        d.ijk(cell, i, j, k);
        float wg, wdep;
        if (!arc_cell_weights(p, i, j, k, f_l[cell], wg, wdep)) continue;
        double ss, e;
        two_sum(s, (double)(wg * wdep), ss, e);
        s = ss;
        c += e;
    }
    block_reduce_pair(s, c, sh_s, sh_c);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = s;
        partials[2 * blockIdx.x + 1] = c;
    }
}

/* One block: fixed-order combination of the block results; writes the fp32 total and the
 * energy per unit weight (reference kernels.py:493-496). */
__global__ void __launch_bounds__(kBlock) k_arc_total(const double *__restrict__ partials,
                                                      int nparts, float Q_w, float eta, float dt,
                                                      float dx3, float *__restrict__ out2,
                                                      float *__restrict__ norm_out)
{
    __shared__ double sh_s[kBlock], sh_c[kBlock];
    double s = 0.0, c = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        double ss, e;
        two_sum(s, partials[2 * b], ss, e);
        s = ss;
        c += partials[2 * b + 1] + e;
    }
    block_reduce_pair(s, c, sh_s, sh_c);
    if (threadIdx.x == 0) {
        float w_total = (float)(s + c);
        float epw = 0.0f;
        if (w_total > 1e-9f) epw = Q_w * eta * dt / (w_total * dx3);
        out2[0] = w_total;
        out2[1] = epw;
        if (norm_out) norm_out[0] = w_total;
    }
}

__global__ void k_arc_apply(float *__restrict__ H, const uint8_t *__restrict__ flags,
                            const float *__restrict__ f_l, ArcParams p, Dims d,
                            const float *__restrict__ tot)
{
    size_t N = d.n();
    size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= N) return;
    float epw = tot[1];
    if (epw <= 0.0f || flags[cell] == GAS) return;
    int i, j, k;
    d.ijk(cell, i, j, k);
    float wg, wdep;
    if (!arc_cell_weights(p, i, j, k, f_l[cell], wg, wdep)) return;
    H[cell] += epw * wg * wdep;
}

static int arc_run(const char *name, float *H, const uint8_t *flags, const float *f_l,
                   float *norm_out, const ArcParams &p, float Q_w, float dt, float dx3, float eta,
                   int nx, int ny, int nz, int layout, void *scratch, void *stream)
{
    WAAM_RETURN_IF(check_dims(name, nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs(name, {H, flags, f_l, scratch}));
    Dims d{nx, ny, nz, layout};
    hipStream_t s = as_stream(stream);
    double *partials = reinterpret_cast<double *>((char *)scratch + kScratchPartials);
    float *tot = reinterpret_cast<float *>((char *)scratch + kScratchTelemetry);
    hipLaunchKernelGGL(k_arc_partials, dim3(kReduceBlocks), dim3(kBlock), 0, s, flags, f_l, p, d,
                       partials);
    hipLaunchKernelGGL(k_arc_total, dim3(1), dim3(kBlock), 0, s, partials, kReduceBlocks, Q_w,
                       eta, dt, dx3, tot, norm_out);
    hipLaunchKernelGGL(k_arc_apply, dim3(grid_for(d.n())), dim3(kBlock), 0, s, H, flags, f_l, p, d,
                       tot);
    return check_launch(name);
}

/* ---------------------------------------------------------------------------------------------
 * Enthalpy clamps and phase map (pointwise).
 * ------------------------------------------------------------------------------------------- */
__global__ void k_ceiling(float *__restrict__ H, const uint8_t *__restrict__ flags, float H_liq,
                          float h_cap_liquid, float h_cap_solid, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N || flags[c] == GAS) return;
    H[c] = ceiling_clamp(H[c], H_liq, h_cap_liquid, h_cap_solid);
}

__global__ void k_floor(float *__restrict__ H, const uint8_t *__restrict__ flags, float h_min,
                        size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N || flags[c] == GAS) return;
    if (H[c] < h_min) H[c] = h_min;
}

__global__ void k_phase(const float *__restrict__ H, float *__restrict__ T, float *__restrict__ f_l,
                        PhaseConsts pc, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float t, fl;
    phase_of(H[c], pc, t, fl);
    T[c] = t;
    f_l[c] = fl;
}

__global__ void k_T_max(const float *__restrict__ T, float *__restrict__ T_max,
                        const uint8_t *__restrict__ flags, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N || flags[c] == GAS) return;
    float t = T[c];
    if (t > T_max[c]) T_max[c] = t;
}

__global__ void k_cooling(const float *__restrict__ T, float *__restrict__ T_prev,
                          float *__restrict__ dT_dt, const uint8_t *__restrict__ flags,
                          float inv_dt, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N || flags[c] == GAS) return;
    float t = T[c];
    dT_dt[c] = (t - T_prev[c]) * inv_dt;
    T_prev[c] = t;
}

__global__ void k_time_above(const float *__restrict__ T, const uint8_t *__restrict__ flags,
                             float *__restrict__ a800, float *__restrict__ a1100,
                             float *__restrict__ asol, float dt, float T800, float T1100,
                             float Ts, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N || flags[c] == GAS) return;
    float t = T[c];
    if (t >= T800) a800[c] += dt;
    if (t >= T1100) a1100[c] += dt;
    if (t >= Ts) asol[c] += dt;
}

/* ---------------------------------------------------------------------------------------------
 * Advection-diffusion (reference kernels.py:1611-1682) in Jacobi form: H_out = H_in + dH(H_in).
 * ------------------------------------------------------------------------------------------- */
__global__ void k_advect_diffuse(const float *__restrict__ H_in, float *__restrict__ H_out,
                                 const float *__restrict__ T, const float *__restrict__ ux,
                                 const float *__restrict__ uy, const float *__restrict__ uz,
                                 const uint8_t *__restrict__ flags, float alpha_lu, float dt,
                                 float cp_rho, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    size_t c = d.idx(i, j, k);
    uint8_t flag = flags[c];
    if (flag == GAS) {
        H_out[c] = H_in[c];
        return;
    }
    H_out[c] = H_in[c] + thermal_dH(H_in, T, ux, uy, uz, flags, i, j, k, flag, alpha_lu, dt,
                                    cp_rho, d);
}

/* apply_thermal_boundary_losses — reference kernels.py:2030-2082 */
__global__ void k_losses(float *__restrict__ H, const float *__restrict__ T,
                         const uint8_t *__restrict__ flags, LossConsts lc, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    size_t c = d.idx(i, j, k);
    if (flags[c] == GAS) return;
    int n_exp = exposed_faces(flags, i, j, k, d);
    if (n_exp == 0) return;
    H[c] -= loss_flux(T[c], lc) * (float)n_exp * lc.dt / lc.dx;
}

/* surface_height_at — reference kernels.py:625-641; max qualifying k, one block */
__global__ void k_surface_height(const float *__restrict__ phi, const uint8_t *__restrict__ flags,
                                 float *__restrict__ out, int i0, int j0, int k_fallback, Dims d)
{
    __shared__ int sh[kBlock];
    int best = -1;
    for (int k = threadIdx.x; k < d.nz; k += blockDim.x) {
        size_t c = d.idx(i0, j0, k);
        if (flags[c] != GAS && phi[c] > 0.05f) best = k > best ? k : best;
    }
    sh[threadIdx.x] = best;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] = imax(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = (float)(sh[0] >= 0 ? sh[0] : k_fallback);
}

}  // namespace waam

using namespace waam;

extern "C" {

int waam_inject_arc_heat(float *H, const uint8_t *flags, const float *f_l, float *norm_out,
                         float arc_i, float arc_j, float arc_k, float Q_w, float sigma, float dt,
                         float dx3, float eta, float penetration_cells, int enable_surface_weight,
                         int nx, int ny, int nz, int layout, void *scratch, void *stream)
{
    ArcParams p{};
    p.kind = 0;
    p.arc_i = arc_i;
    p.arc_j = arc_j;
    p.arc_k = arc_k;
    p.inv2s2 = 1.0f / (2.0f * sigma * sigma);
    p.pen = penetration_cells;
    p.enable = enable_surface_weight;
    return arc_run("waam_inject_arc_heat", H, flags, f_l, norm_out, p, Q_w, dt, dx3, eta, nx, ny,
                   nz, layout, scratch, stream);
}

int waam_inject_goldak_heat(float *H, const uint8_t *flags, const float *f_l, float *norm_out,
                            float arc_i, float arc_j, float arc_k, float Q_w, float dt,
                            float dx3, float eta, float travel_sign, float ff, float fr,
                            float a_front, float a_rear, float b_axis, float c_axis,
                            float penetration_cells, int enable_surface_weight, int nx, int ny,
                            int nz, int layout, void *scratch, void *stream)
{
    ArcParams p{};
    p.kind = 1;
    p.arc_i = arc_i;
    p.arc_j = arc_j;
    p.arc_k = arc_k;
    p.sgn = travel_sign;
    p.ff = ff;
    p.fr = fr;
    p.af = a_front;
    p.ar = a_rear;
    p.b = b_axis;
    p.cax = c_axis;
    p.pen = penetration_cells;
    p.enable = enable_surface_weight;
    return arc_run("waam_inject_goldak_heat", H, flags, f_l, norm_out, p, Q_w, dt, dx3, eta, nx,
                   ny, nz, layout, scratch, stream);
}

int waam_surface_height_at(const float *phi, const uint8_t *flags, float *out, int i0, int j0,
                           int nz_fallback, int nx, int ny, int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_surface_height_at", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_surface_height_at", {phi, flags, out}));
    if (i0 < 0 || i0 >= nx || j0 < 0 || j0 >= ny)
        return set_error(WAAM_ERR_INVALID, "waam_surface_height_at: column (%d,%d) outside grid",
                         i0, j0);
    int kf = nz_fallback - 1 > 1 ? nz_fallback - 1 : 1;
    hipLaunchKernelGGL(k_surface_height, dim3(1), dim3(kBlock), 0, as_stream(stream), phi, flags,
                       out, i0, j0, kf, Dims{nx, ny, nz, layout});
    return check_launch("waam_surface_height_at");
}

int waam_clamp_enthalpy_ceiling_scalar(float *H, const uint8_t *flags, float cp_rho,
                                       float T_solidus, float T_liquidus, float T_vapor_cap,
                                       float L_rho, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_clamp_enthalpy_ceiling_scalar: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_clamp_enthalpy_ceiling_scalar", {H, flags}));
    float H_sol = cp_rho * T_solidus;
    float H_liq = H_sol + L_rho;
    float cap_l = H_liq + cp_rho * (T_vapor_cap - T_liquidus);
    float cap_s = cp_rho * T_vapor_cap;
    hipLaunchKernelGGL(k_ceiling, dim3(grid_for((size_t)n)), dim3(kBlock), 0, as_stream(stream), H,
                       flags, H_liq, cap_l, cap_s, (size_t)n);
    return check_launch("waam_clamp_enthalpy_ceiling_scalar");
}

int waam_clamp_enthalpy_floor_scalar(float *H, float cp_rho, const uint8_t *flags,
                                     float T_floor, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_clamp_enthalpy_floor_scalar: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_clamp_enthalpy_floor_scalar", {H, flags}));
    hipLaunchKernelGGL(k_floor, dim3(grid_for((size_t)n)), dim3(kBlock), 0, as_stream(stream), H,
                       flags, cp_rho * T_floor, (size_t)n);
    return check_launch("waam_clamp_enthalpy_floor_scalar");
}

int waam_advect_diffuse_temperature(const float *H_in, float *H_out, const float *T,
                                    const float *ux, const float *uy, const float *uz,
                                    const uint8_t *flags, float alpha_lu, float dt,
                                    float cp_rho, int nx, int ny, int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_advect_diffuse_temperature", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_advect_diffuse_temperature", {H_in, H_out, T, ux, uy, uz, flags}));
    if (H_in == H_out)
        return set_error(WAAM_ERR_INVALID, "waam_advect_diffuse_temperature: H_in aliases H_out");
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_advect_diffuse, grid3(d), block3(), 0, as_stream(stream), H_in, H_out, T,
                       ux, uy, uz, flags, alpha_lu, dt, cp_rho, d);
    return check_launch("waam_advect_diffuse_temperature");
}

int waam_apply_thermal_boundary_losses(float *H, const float *T, const uint8_t *flags,
                                       float T_amb, float h_conv, float eps_rad,
                                       int enable_conv, int enable_rad, float cp_rho, float dt,
                                       float dx, float sigma_sb, int nx, int ny, int nz, int layout,
                                       void *stream)
{
    (void)cp_rho;
    WAAM_RETURN_IF(check_dims("waam_apply_thermal_boundary_losses", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_apply_thermal_boundary_losses", {H, T, flags}));
    Dims d{nx, ny, nz, layout};
    LossConsts lc{T_amb, h_conv, eps_rad, sigma_sb, dt, dx, enable_conv, enable_rad};
    hipLaunchKernelGGL(k_losses, grid3(d), block3(), 0, as_stream(stream), H, T, flags, lc, d);
    return check_launch("waam_apply_thermal_boundary_losses");
}

int waam_update_phase(const float *H, float *T, float *f_l, float cp_rho, float L_rho,
                      float T_solidus, float T_liquidus, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_update_phase: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_update_phase", {H, T, f_l}));
    PhaseConsts pc = make_phase_consts(cp_rho, L_rho, T_solidus, T_liquidus);
    hipLaunchKernelGGL(k_phase, dim3(grid_for((size_t)n)), dim3(kBlock), 0, as_stream(stream), H,
                       T, f_l, pc, (size_t)n);
    return check_launch("waam_update_phase");
}

int waam_update_T_max(const float *T, float *T_max, const uint8_t *flags, int64_t n,
                      void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_update_T_max: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_update_T_max", {T, T_max, flags}));
    hipLaunchKernelGGL(k_T_max, dim3(grid_for((size_t)n)), dim3(kBlock), 0, as_stream(stream), T,
                       T_max, flags, (size_t)n);
    return check_launch("waam_update_T_max");
}

int waam_update_cooling_rate(const float *T, float *T_prev, float *dT_dt, const uint8_t *flags,
                             float dt_phys, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_update_cooling_rate: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_update_cooling_rate", {T, T_prev, dT_dt, flags}));
    float inv_dt = 1.0f / (dt_phys + 1e-12f);
    hipLaunchKernelGGL(k_cooling, dim3(grid_for((size_t)n)), dim3(kBlock), 0, as_stream(stream), T,
                       T_prev, dT_dt, flags, inv_dt, (size_t)n);
    return check_launch("waam_update_cooling_rate");
}

int waam_update_time_above_T(const float *T, const uint8_t *flags, float *t800, float *t1100,
                             float *tsol, float dt, float T_800, float T_1100, float T_solidus,
                             int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_update_time_above_T: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_update_time_above_T", {T, flags, t800, t1100, tsol}));
    hipLaunchKernelGGL(k_time_above, dim3(grid_for((size_t)n)), dim3(kBlock), 0, as_stream(stream),
                       T, flags, t800, t1100, tsol, dt, T_800, T_1100, T_solidus, (size_t)n);
    return check_launch("waam_update_time_above_T");
}

}  // extern "C"
