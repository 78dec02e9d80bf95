// lbm.hip — D3Q19 lattice Boltzmann kernels: initialisation, two-rate central-moment (MRT)
// and BGK (SRT) collisions with Guo forcing and Carman-Kozeny drag, the pull stream with
// bounce-back, the Mach clamp and the droplet-impact equilibrium reset.
//
// Populations are SoA: f[q*N + c]. One thread owns one cell; consecutive threads own
// consecutive k, so every population plane is read and written fully coalesced.
#include "lbm_cell.h"

namespace waam {

/* init_grid — reference kernels.py:62-120 */
__global__ void k_init_grid(float *__restrict__ f, float *__restrict__ rho, float *__restrict__ ux,
                            float *__restrict__ uy, float *__restrict__ uz, float *__restrict__ T,
                            float *__restrict__ H, float *__restrict__ f_l,
                            float *__restrict__ phi, uint8_t *__restrict__ flags, float T_amb,
                            float rho0, float cp_rho, int nz_solid, int pi0, int pi1, int pj0,
                            int pj1, Dims d)
{
    size_t N = d.n();
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N;
         c += (size_t)gridDim.x * blockDim.x) {
        int i, j, k;
        d.ijk(c, i, j, k);
#pragma unroll
        for (int q = 0; q < Q; ++q) f[(size_t)q * N + c] = wq(q) * rho0;
        rho[c] = rho0;
        ux[c] = 0.0f;
        uy[c] = 0.0f;
        uz[c] = 0.0f;
        T[c] = T_amb;
        H[c] = cp_rho * T_amb;
        if (nz_solid < 0) {
            flags[c] = FLUID;
            phi[c] = 1.0f;
            f_l[c] = 1.0f;
        } else {
            f_l[c] = 0.0f;
            bool in_plate = i >= pi0 && i < pi1 && j >= pj0 && j < pj1;
            if (in_plate && k < nz_solid) {
                phi[c] = 1.0f;
                flags[c] = SOLID;
            } else {
                phi[c] = 0.0f;
                flags[c] = GAS;
            }
        }
    }
}

template <int kMode>  // 0 = MRT, 1 = SRT (omega given), 2 = SRT (omega from tau_field)
__global__ void k_collide(const float *__restrict__ f_src, float *__restrict__ f_dst,
                          float *__restrict__ rho, float *__restrict__ ux, float *__restrict__ uy,
                          float *__restrict__ uz, const float *__restrict__ Fx,
                          const float *__restrict__ Fy, const float *__restrict__ Fz,
                          const float *__restrict__ f_l, const uint8_t *__restrict__ flags,
                          const float *__restrict__ tau_field, float omega_s, float omega_b,
                          float C_darcy, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    uint8_t fl = flags[c];
    if (fl == SOLID || fl == GAS) return;
    float f[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = f_src[(size_t)q * N + c];
    CellOut o;
    if (kMode == 0) {
        mrt_cell(f, Fx[c], Fy[c], Fz[c], f_l[c], omega_s, omega_b, C_darcy, o);
    } else {
        float om = omega_s;
        if (kMode == 2) om = 1.0f / (tau_field[c] + 1e-6f);
        srt_cell(f, Fx[c], Fy[c], Fz[c], f_l[c], om, C_darcy, o);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) f_dst[(size_t)q * N + c] = o.f[q];
    rho[c] = o.rho;
    ux[c] = o.u;
    uy[c] = o.v;
    uz[c] = o.w;
}

/* stream — reference kernels.py:2971-3010 */
__global__ void k_stream(const float *__restrict__ f_src, float *__restrict__ f_dst,
                         const uint8_t *__restrict__ flags, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    size_t N = d.n();
    size_t c = d.idx(i, j, k);
    uint8_t flag = flags[c];
    if (flag == SOLID) return;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        int ni = i - ex(q), nj = j - ey(q), nk = k - ez(q);
        ni = ni < 0 ? ni + d.nx : (ni >= d.nx ? ni - d.nx : ni);
        nj = imax(0, imin(nj, d.ny - 1));
        nk = imax(0, imin(nk, d.nz - 1));
        size_t n = d.idx(ni, nj, nk);
        float v;
        if (flag == GAS) v = f_src[(size_t)q * N + n];
        else if (flags[n] == SOLID) v = f_src[(size_t)opp(q) * N + c];
        else v = f_src[(size_t)q * N + n];
        f_dst[(size_t)q * N + c] = v;
    }
}

/* clamp_velocity_mach — reference kernels.py:3312-3335 */
__global__ void k_clamp_velocity(float *__restrict__ ux, float *__restrict__ uy,
                                 float *__restrict__ uz, const uint8_t *__restrict__ flags,
                                 float u_max, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    uint8_t fl = flags[c];
    if (fl == SOLID || fl == GAS) return;
    float vx = ux[c], vy = uy[c], vz = uz[c];
    float mag = sqrtf(vx * vx + vy * vy + vz * vz);
    if (mag > u_max) {
        float s = u_max / (mag + 1e-12f);
        ux[c] = vx * s;
        uy[c] = vy * s;
        uz[c] = vz * s;
    }
}

/* feed_wire_momentum_impact — reference kernels.py:1176-1233; the sub-box is launched exactly */
__global__ void k_momentum_impact(float *__restrict__ f_src, const float *__restrict__ rho,
                                  float *__restrict__ ux, float *__restrict__ uy,
                                  float *__restrict__ uz, const uint8_t *__restrict__ flags,
                                  const float *__restrict__ f_l, float arc_i, float arc_j,
                                  float arc_k, float dr, float vx, float vy, float vz, int i0,
                                  int j0, int k0, int ni, int nj, int nk, Dims d,
                                  const float *__restrict__ arc_k_dev)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni * nj * nk) return;
    int k = k0 + t % nk;
    int j = j0 + (t / nk) % nj;
    int i = i0 + t / (nk * nj);
    if (arc_k_dev) {
        /* arc_k resolved on the device: the launch covers every k, and the host's k range
         * [(int)arc_k, (int)(arc_k + 3 dr) + 1) is applied here */
        arc_k = *arc_k_dev;
        const int k_min = imax((int)arc_k, 0);
        const int k_max = imin((int)(arc_k + dr * 3.0f) + 1, d.nz);
        if (k < k_min || k >= k_max) return;
    }
    size_t N = d.n();
    size_t c = d.idx(i, j, k);
    if (flags[c] == GAS) return;
    if (f_l[c] < 0.45f) return;
    float drop_cz = arc_k + dr + 1.0f;
    float di = (float)i - arc_i, dj = (float)j - arc_j, dk = (float)k - drop_cz;
    float r2 = di * di + dj * dj + dk * dk;
    float rij2 = di * di + dj * dj;
    bool in_drop = r2 <= dr * dr || (rij2 <= dr * dr && (float)k >= arc_k + 1.0f);
    if (!in_drop) return;
    float u2 = vx * vx + vy * vy + vz * vz;
    float r_loc = rho[c];
    ux[c] = vx;
    uy[c] = vy;
    uz[c] = vz;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        float eu = (float)ex(q) * vx + (float)ey(q) * vy + (float)ez(q) * vz;
        f_src[(size_t)q * N + c] = wq(q) * r_loc * (1.0f + 3.0f * eu + 4.5f * eu * eu - 1.5f * u2);
    }
}

}  // namespace waam

using namespace waam;

extern "C" {

int waam_init_grid(float *f, float *rho, float *ux, float *uy, float *uz, float *T, float *H,
                   float *f_l, float *phi, uint8_t *flags, float T_ambient, float rho0,
                   float cp_rho, int nz_solid, int plate_i0, int plate_i1, int plate_j0,
                   int plate_j1, int nx, int ny, int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_init_grid", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_init_grid", {f, rho, ux, uy, uz, T, H, f_l, phi, flags}));
    Dims d{nx, ny, nz, layout};
    unsigned g = grid_for(d.n());
    if (g > 65536u) g = 65536u;
    hipLaunchKernelGGL(k_init_grid, dim3(g), dim3(kBlock), 0, as_stream(stream), f, rho, ux, uy,
                       uz, T, H, f_l, phi, flags, T_ambient, rho0, cp_rho, nz_solid, plate_i0,
                       plate_i1, plate_j0, plate_j1, d);
    return check_launch("waam_init_grid");
}

static int collide_common(int mode, const char *name, const float *f_src, float *f_dst,
                          float *rho, float *ux, float *uy, float *uz, const float *Fx,
                          const float *Fy, const float *Fz, const float *f_l,
                          const uint8_t *flags, const float *tau_field, float om_s, float om_b,
                          float C, int nx, int ny, int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims(name, nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs(name, {f_src, f_dst, rho, ux, uy, uz, Fx, Fy, Fz, f_l, flags}));
    if (f_src == f_dst) return set_error(WAAM_ERR_INVALID, "%s: f_src aliases f_dst", name);
    Dims d{nx, ny, nz, layout};
    size_t N = d.n();
    hipStream_t s = as_stream(stream);
    if (mode == 0)
        hipLaunchKernelGGL(k_collide<0>, dim3(grid_for(N)), dim3(kBlock), 0, s, f_src, f_dst, rho,
                           ux, uy, uz, Fx, Fy, Fz, f_l, flags, tau_field, om_s, om_b, C, N);
    else if (mode == 1)
        hipLaunchKernelGGL(k_collide<1>, dim3(grid_for(N)), dim3(kBlock), 0, s, f_src, f_dst, rho,
                           ux, uy, uz, Fx, Fy, Fz, f_l, flags, tau_field, om_s, om_b, C, N);
    else
        hipLaunchKernelGGL(k_collide<2>, dim3(grid_for(N)), dim3(kBlock), 0, s, f_src, f_dst, rho,
                           ux, uy, uz, Fx, Fy, Fz, f_l, flags, tau_field, om_s, om_b, C, N);
    return check_launch(name);
}

int waam_collide_mrt(const float *f_src, float *f_dst, float *rho, float *ux, float *uy,
                     float *uz, const float *Fx, const float *Fy, const float *Fz,
                     const float *f_l, const uint8_t *flags, float omega_s, float omega_b,
                     float C_darcy, int nx, int ny, int nz, int layout, void *stream)
{
    return collide_common(0, "waam_collide_mrt", f_src, f_dst, rho, ux, uy, uz, Fx, Fy, Fz, f_l,
                          flags, nullptr, omega_s, omega_b, C_darcy, nx, ny, nz, layout, stream);
}

int waam_collide_srt(const float *f_src, float *f_dst, float *rho, float *ux, float *uy,
                     float *uz, const float *Fx, const float *Fy, const float *Fz,
                     const float *f_l, const uint8_t *flags, float tau, float omega,
                     float dt_lu, float C_darcy, int nx, int ny, int nz, int layout, void *stream)
{
    (void)tau;
    (void)dt_lu;
    return collide_common(1, "waam_collide_srt", f_src, f_dst, rho, ux, uy, uz, Fx, Fy, Fz, f_l,
                          flags, nullptr, omega, omega, C_darcy, nx, ny, nz, layout, stream);
}

int waam_collide_srt_variable_tau(const float *f_src, float *f_dst, float *rho, float *ux,
                                  float *uy, float *uz, const float *Fx, const float *Fy,
                                  const float *Fz, const float *f_l, const uint8_t *flags,
                                  const float *tau_field, float C_darcy, int nx, int ny, int nz, int layout,
                                  void *stream)
{
    if (!tau_field)
        return set_error(WAAM_ERR_INVALID, "waam_collide_srt_variable_tau: tau_field is NULL");
    return collide_common(2, "waam_collide_srt_variable_tau", f_src, f_dst, rho, ux, uy, uz, Fx,
                          Fy, Fz, f_l, flags, tau_field, 0.0f, 0.0f, C_darcy, nx, ny, nz,
                          layout, stream);
}

int waam_stream(const float *f_src, float *f_dst, const uint8_t *flags, int nx, int ny, int nz, int layout,
                void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_stream", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_stream", {f_src, f_dst, flags}));
    if (f_src == f_dst) return set_error(WAAM_ERR_INVALID, "waam_stream: f_src aliases f_dst");
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_stream, grid3(d), block3(), 0, as_stream(stream), f_src, f_dst, flags, d);
    return check_launch("waam_stream");
}

int waam_clamp_velocity_mach(float *ux, float *uy, float *uz, const uint8_t *flags,
                             float u_max_lu, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_clamp_velocity_mach: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_clamp_velocity_mach", {ux, uy, uz, flags}));
    hipLaunchKernelGGL(k_clamp_velocity, dim3(grid_for((size_t)n)), dim3(kBlock), 0,
                       as_stream(stream), ux, uy, uz, flags, u_max_lu, (size_t)n);
    return check_launch("waam_clamp_velocity_mach");
}

int waam_feed_wire_momentum_impact(float *f_src, const float *rho, float *ux, float *uy,
                                   float *uz, const uint8_t *flags, const float *f_l,
                                   float arc_i, float arc_j, float arc_k, float droplet_radius,
                                   float vx_lu, float vy_lu, float vz_lu, int nx, int ny, int nz, int layout,
                                   void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_feed_wire_momentum_impact", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_feed_wire_momentum_impact", {f_src, rho, ux, uy, uz, flags, f_l}));
    const float dr = droplet_radius;
    int i_min = (int)(arc_i - dr * 2.5f); i_min = i_min > 0 ? i_min : 0;
    int i_max = (int)(arc_i + dr * 2.5f) + 1; i_max = i_max < nx ? i_max : nx;
    int j_min = (int)(arc_j - dr * 2.5f); j_min = j_min > 0 ? j_min : 0;
    int j_max = (int)(arc_j + dr * 2.5f) + 1; j_max = j_max < ny ? j_max : ny;
    int k_min = (int)(arc_k); k_min = k_min > 0 ? k_min : 0;
    int k_max = (int)(arc_k + dr * 3.0f) + 1; k_max = k_max < nz ? k_max : nz;
    int ni = i_max - i_min, nj = j_max - j_min, nk = k_max - k_min;
    if (ni <= 0 || nj <= 0 || nk <= 0) return WAAM_OK;
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_momentum_impact, dim3(grid_for((size_t)ni * nj * nk)), dim3(kBlock), 0,
                       as_stream(stream), f_src, rho, ux, uy, uz, flags, f_l, arc_i, arc_j,
                       arc_k, dr, vx_lu, vy_lu, vz_lu, i_min, j_min, k_min, ni, nj, nk, d,
                       (const float *)nullptr);
    return check_launch("waam_feed_wire_momentum_impact");
}

}  // extern "C"

namespace waam {

int launch_momentum_impact_dev(float *f_src, const float *rho, float *ux, float *uy, float *uz,
                               const uint8_t *flags, const float *f_l, float arc_i, float arc_j,
                               const float *arc_k_dev, float dr, float vx, float vy, float vz,
                               Dims d, hipStream_t st)
{
    int i_min = (int)(arc_i - dr * 2.5f); i_min = i_min > 0 ? i_min : 0;
    int i_max = (int)(arc_i + dr * 2.5f) + 1; i_max = i_max < d.nx ? i_max : d.nx;
    int j_min = (int)(arc_j - dr * 2.5f); j_min = j_min > 0 ? j_min : 0;
    int j_max = (int)(arc_j + dr * 2.5f) + 1; j_max = j_max < d.ny ? j_max : d.ny;
    const int ni = i_max - i_min, nj = j_max - j_min, nk = d.nz;
    if (ni <= 0 || nj <= 0) return WAAM_OK;
    hipLaunchKernelGGL(k_momentum_impact, dim3(grid_for((size_t)ni * nj * nk)), dim3(kBlock), 0,
                       st, f_src, rho, ux, uy, uz, flags, f_l, arc_i, arc_j, 0.0f, dr, vx, vy, vz,
                       i_min, j_min, 0, ni, nj, nk, d, arc_k_dev);
    return check_launch("waam_step_fused_droplet (momentum impact)");
}

}  // namespace waam
