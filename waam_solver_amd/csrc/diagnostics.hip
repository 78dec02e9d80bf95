// diagnostics.hip — export-side fields and reductions: vorticity magnitude (reference
// kernels.py:3380-3419, compute_vorticity_magnitude) and the max-norm of a vector field (the
// reference physics/force_diagnostics.py:19-24 reads the three force fields back to the host
// for it; here one device reduction leaves one float).
#include "waam_common.h"

namespace waam {

__global__ void k_vorticity(const float *__restrict__ ux, const float *__restrict__ uy,
                            const float *__restrict__ uz, const uint8_t *__restrict__ flags,
                            float *__restrict__ out, float dt, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    size_t c = d.idx(i, j, k);
    if (flags[c] == GAS) {
        out[c] = 0.0f;
        return;
    }
    float scale = 1.0f / (dt + 1e-12f);
    int ip = imin(i + 1, d.nx - 1), im = imax(i - 1, 0);
    int jp = imin(j + 1, d.ny - 1), jm = imax(j - 1, 0);
    int kp = imin(k + 1, d.nz - 1), km = imax(k - 1, 0);
    float duz_dy = 0.5f * (uz[d.idx(i, jp, k)] - uz[d.idx(i, jm, k)]);
    float duy_dz = 0.5f * (uy[d.idx(i, j, kp)] - uy[d.idx(i, j, km)]);
    float dux_dz = 0.5f * (ux[d.idx(i, j, kp)] - ux[d.idx(i, j, km)]);
    float duz_dx = 0.5f * (uz[d.idx(ip, j, k)] - uz[d.idx(im, j, k)]);
    float duy_dx = 0.5f * (uy[d.idx(ip, j, k)] - uy[d.idx(im, j, k)]);
    float dux_dy = 0.5f * (ux[d.idx(i, jp, k)] - ux[d.idx(i, jm, k)]);
    float wx = (duz_dy - duy_dz) * scale;
    float wy = (dux_dz - duz_dx) * scale;
    float wz = (duy_dx - dux_dy) * scale;
    out[c] = sqrtf(wx * wx + wy * wy + wz * wz);
}

__global__ void k_zero_u32(unsigned *out) { out[0] = 0u; }

/* out = max over cells of sqrt(fx^2 + fy^2 + fz^2); a non-negative float orders like its bits,
 * so an unsigned atomic max is exact and order-free. A NaN wins every comparison here (its bits
 * exceed those of +inf), so a NaN force shows as a NaN maximum, as numpy's max reports it. */
__device__ __forceinline__ float max_nan(float a, float b) { return (b > a || b != b) ? b : a; }

__global__ void __launch_bounds__(kBlock) k_max_norm(const float *__restrict__ fx,
                                                     const float *__restrict__ fy,
                                                     const float *__restrict__ fz,
                                                     unsigned *out, size_t N)
{
    float best = 0.0f;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N;
         c += (size_t)gridDim.x * blockDim.x) {
        float a = fx[c], b = fy[c], e = fz[c];
        float m = sqrtf(a * a + b * b + e * e);
        best = max_nan(best, m);
    }
    for (int off = 32; off > 0; off >>= 1) {
        best = max_nan(best, __shfl_xor(best, off));
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(best));
}

/* probe sample: out[2p] = T[cell_p], out[2p + 1] = f_l[cell_p]; a negative cell (a probe outside
 * the window) gives NaN in both */
__global__ void k_probe_gather(const float *__restrict__ T, const float *__restrict__ fl,
                               const int64_t *__restrict__ cells, int n_probes, size_t N,
                               float *__restrict__ out)
{
    int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= n_probes) return;
    int64_t c = cells[p];
    bool in = c >= 0 && (size_t)c < N;
    out[2 * p] = in ? T[c] : __int_as_float(0x7fc00000);
    out[2 * p + 1] = in ? fl[c] : __int_as_float(0x7fc00000);
}

}  // namespace waam

using namespace waam;

extern "C" {

int waam_compute_vorticity_magnitude(const float *ux, const float *uy, const float *uz,
                                     const uint8_t *flags, float *vort_out, float dx, float dt,
                                     int nx, int ny, int nz, int layout, void *stream)
{
    (void)dx;  // the lattice-to-physical factor is 1/dt (reference kernels.py:3393-3398)
    WAAM_RETURN_IF(check_dims("waam_compute_vorticity_magnitude", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_compute_vorticity_magnitude", {ux, uy, uz, flags, vort_out}));
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_vorticity, grid3(d), block3(), 0, as_stream(stream), ux, uy, uz, flags,
                       vort_out, dt, d);
    return check_launch("waam_compute_vorticity_magnitude");
}

int waam_max_norm3(const float *fx, const float *fy, const float *fz, float *out, int64_t n,
                   void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_max_norm3: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_max_norm3", {fx, fy, fz, out}));
    hipStream_t s = as_stream(stream);
    unsigned *o = reinterpret_cast<unsigned *>(out);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, o);
    unsigned g = grid_for((size_t)n);
    if (g > 4096u) g = 4096u;
    hipLaunchKernelGGL(k_max_norm, dim3(g), dim3(kBlock), 0, s, fx, fy, fz, o, (size_t)n);
    return check_launch("waam_max_norm3");
}

int waam_probe_gather(const float *T, const float *f_l, const int64_t *cells, int n_probes,
                      int64_t n_cells, float *out, void *stream)
{
    if (n_probes <= 0 || n_cells <= 0)
        return set_error(WAAM_ERR_INVALID, "waam_probe_gather: n_probes %d, n_cells %lld",
                         n_probes, (long long)n_cells);
    WAAM_RETURN_IF(check_ptrs("waam_probe_gather", {T, f_l, cells, out}));
    hipLaunchKernelGGL(k_probe_gather, dim3((unsigned)((n_probes + 63) / 64)), dim3(64), 0,
                       as_stream(stream), T, f_l, cells, n_probes, (size_t)n_cells, out);
    return check_launch("waam_probe_gather");
}

}  // extern "C"
