// window.hip — moving simulation window (reference kernels.py:2620-2723,
// shift_simulation_window_x).
//
// The reference runs one serial loop per (j, k) row in ascending i, so every field becomes a
// pure shift, new[i] = old[i + n] for i < nx - n, followed by a reset of the last n columns to
// the ambient plate/gas state. Here one workgroup owns one row and walks it in ascending chunks
// of its width: a chunk reads old[i + n] (never yet written: every earlier chunk wrote only
// positions below it), synchronises, then writes; the reset of the strip follows the last
// chunk. The row is contiguous in the (nz, ny, nx) storage and strided in (nx, ny, nz).
#include "waam_common.h"

namespace waam {

constexpr int kShiftFloat = 22;  // scalar fp32 fields, in the reference's argument order

struct ShiftFields {
    float *fl[kShiftFloat];  // T H f_l phi T_max T_prev dT_dt t800 t1100 tsol rho ux uy uz
                             // Fx Fy Fz cp_rho alpha_lu dgamma_lu tau (+1 unused)
    uint8_t *flags;
    float *f_a, *f_b;
};

template <class V>
__device__ __forceinline__ void shift_row(V *a, int n, int j, int k, const Dims &d)
{
    for (int base = 0; base < d.nx - n; base += blockDim.x) {
        int i = base + (int)threadIdx.x;
        bool ok = i < d.nx - n;
        V v{};
        if (ok) v = a[d.idx(i + n, j, k)];
        __syncthreads();
        if (ok) a[d.idx(i, j, k)] = v;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock) k_shift_window_x(ShiftFields s, int n_fields, int n,
                                                           float T_amb, float cp_rho, float rho0,
                                                           int nz_solid, Dims d)
{
    int row = blockIdx.x;  // j + ny * k
    int j = row % d.ny, k = row / d.ny;
    size_t N = d.n();
    for (int a = 0; a < n_fields; ++a) shift_row(s.fl[a], n, j, k, d);
    shift_row(s.flags, n, j, k, d);
    for (int q = 0; q < Q; ++q) {
        shift_row(s.f_a + (size_t)q * N, n, j, k, d);
        shift_row(s.f_b + (size_t)q * N, n, j, k, d);
    }
    /* the strip i >= nx - n: ambient plate below nz_solid, gas above, rest equilibrium */
    int i0 = d.nx - n < 0 ? 0 : d.nx - n;
    for (int i = i0 + (int)threadIdx.x; i < d.nx; i += blockDim.x) {
        size_t c = d.idx(i, j, k);
        float *const *f = s.fl;
        f[0][c] = T_amb;            // T
        f[1][c] = cp_rho * T_amb;   // H
        f[2][c] = 0.0f;             // f_l
        f[10][c] = rho0;            // rho
        f[11][c] = 0.0f;            // ux
        f[12][c] = 0.0f;            // uy
        f[13][c] = 0.0f;            // uz
        f[14][c] = 0.0f;            // Fx
        f[15][c] = 0.0f;            // Fy
        f[16][c] = 0.0f;            // Fz
        f[5][c] = T_amb;            // T_prev
        f[6][c] = 0.0f;             // dT_dt
        f[4][c] = T_amb;            // T_max
        f[7][c] = 0.0f;             // time above 800 C
        f[8][c] = 0.0f;             // time above 1100 C
        f[9][c] = 0.0f;             // time above solidus
        bool solid = k < nz_solid;
        s.flags[c] = solid ? (uint8_t)SOLID : (uint8_t)GAS;
        f[3][c] = solid ? 1.0f : 0.0f;  // phi
        for (int q = 0; q < Q; ++q) {
            float feq = wq(q) * rho0;
            s.f_a[(size_t)q * N + c] = feq;
            s.f_b[(size_t)q * N + c] = feq;
        }
    }
}

}  // namespace waam

using namespace waam;

extern "C" int waam_shift_simulation_window_x(
    int n_shift, float *f_a, float *f_b, float *T, float *H, float *f_l, float *phi,
    uint8_t *flags, float *T_max, float *T_prev, float *dT_dt, float *time_above_800,
    float *time_above_1100, float *time_above_solidus, float *rho, float *ux, float *uy,
    float *uz, float *Fx, float *Fy, float *Fz, float *cp_rho_field, float *alpha_lu_field,
    float *dgamma_lu_field, float *tau_field, float T_amb, float cp_rho, float rho0, int nz_solid,
    int nx, int ny, int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_shift_simulation_window_x", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_shift_simulation_window_x",
                              {f_a, f_b, T, H, f_l, phi, flags, T_max, T_prev, dT_dt,
                               time_above_800, time_above_1100, time_above_solidus, rho, ux, uy,
                               uz, Fx, Fy, Fz, cp_rho_field, alpha_lu_field, dgamma_lu_field,
                               tau_field}));
    if (n_shift <= 0)
        return set_error(WAAM_ERR_INVALID, "waam_shift_simulation_window_x: n_shift %d <= 0",
                         n_shift);
    if (f_a == f_b)
        return set_error(WAAM_ERR_INVALID, "waam_shift_simulation_window_x: f_a aliases f_b");
    if ((long long)ny * nz > 0x7fffffffLL)
        return set_error(WAAM_ERR_INVALID, "waam_shift_simulation_window_x: too many rows");
    Dims d{nx, ny, nz, layout};
    ShiftFields s{};
    float *list[] = {T,     H,     f_l,  phi,  T_max, T_prev, dT_dt, time_above_800,
                     time_above_1100, time_above_solidus, rho, ux, uy, uz, Fx, Fy, Fz,
                     cp_rho_field, alpha_lu_field, dgamma_lu_field, tau_field};
    int n_fields = (int)(sizeof(list) / sizeof(list[0]));
    for (int a = 0; a < n_fields; ++a) s.fl[a] = list[a];
    s.flags = flags;
    s.f_a = f_a;
    s.f_b = f_b;
    hipLaunchKernelGGL(k_shift_window_x, dim3((unsigned)(ny * nz)), dim3(kBlock), 0,
                       as_stream(stream), s, n_fields, n_shift, T_amb, cp_rho, rho0, nz_solid, d);
    return check_launch("waam_shift_simulation_window_x");
}
