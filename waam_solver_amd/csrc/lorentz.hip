// lorentz.hip — the welding-current Lorentz force of the full physics tier (reference
// kernels.py:1236-1513, driven by physics/weld_forces.py:196-281).
//
// One entry per reference kernel: electrical conductivity, the normalised arc current source, a
// weighted Jacobi sweep for the potential with the plate grounded, the L1 change used for the
// convergence test, J = -sigma grad(phi), the axial current binned by radius per z-slab, its
// radial prefix (enclosed current), the axisymmetric self field B_theta = mu0 I_enc / (2 pi r)
// and the J x B body force. The reference's float atomic sums (source total, L1 norms, radial
// bins) become compensated fp64 sums rounded once to fp32, which does not depend on the order
// of the terms.
#include "arc_cell.h"
#include "waam_common.h"

namespace waam {

constexpr int kLorentzSumBlocks = kReduceBlocks;
constexpr int kBinGroup = 8;  // radial bins per workgroup of the binning kernel

__device__ __forceinline__ float sigma_at(const float *s, int i, int j, int k, const Dims &d)
{
    return s[d.idx(imax(0, imin(d.nx - 1, i)), imax(0, imin(d.ny - 1, j)),
                   imax(0, imin(d.nz - 1, k)))];
}

/* elec_build_sigma kernels.py:1236-1251 */
__global__ void k_elec_build_sigma(float *__restrict__ sigma, const float *__restrict__ f_l,
                                   const uint8_t *__restrict__ flags, float s_liq, float s_sol,
                                   size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float fl = f_l[c];
    uint8_t f = flags[c];
    sigma[c] = f == GAS ? 1e-8f : (fl > 0.55f ? s_liq : s_sol);
}

__global__ void k_fill(float *__restrict__ a, float v, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < N) a[c] = v;
}

struct ElecSource {
    float arc_i, arc_j, arc_k, inv2s2, current, cell_vol;
    int k_min, k_max;
};

/* elec_inject_arc_source kernels.py:1260-1286: cells of the slab range get the Gaussian
 * current density, the others keep their value (0 after elec_clear_source) */
__global__ void k_elec_source(float *__restrict__ source, ElecSource e, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    if (k < e.k_min || k > e.k_max) return;
    float di = (float)i - e.arc_i, dj = (float)j - e.arc_j, dk = (float)k - e.arc_k;
    float r2 = di * di + dj * dj + 0.25f * dk * dk;
    float w = waam_expf(-r2 * e.inv2s2);
    source[d.idx(i, j, k)] = e.current * w / (e.cell_vol + 1e-6f);
}

/* elec_normalize_source kernels.py:1289-1301, first loop: compensated partial sums of
 * source * cell volume */
__global__ void __launch_bounds__(kBlock) k_elec_source_sum(const float *__restrict__ source,
                                                            float cell_vol,
                                                            double *__restrict__ parts, size_t N)
{
    __shared__ double sh_s[kBlock], sh_c[kBlock];
    double sum = 0.0, comp = 0.0;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N;
         c += (size_t)gridDim.x * blockDim.x) {
        double ss, er;
        two_sum(sum, (double)(source[c] * cell_vol), ss, er);
        sum = ss;
        comp += er;
    }
    block_reduce_pair(sum, comp, sh_s, sh_c);
    if (threadIdx.x == 0) {
        parts[2 * blockIdx.x] = sum;
        parts[2 * blockIdx.x + 1] = comp;
    }
}

/* fixed-order total of nb partial pairs, rounded to fp32 into out[0] */
__global__ void __launch_bounds__(kBlock) k_pairs_total(const double *__restrict__ parts, int nb,
                                                        float *__restrict__ out)
{
    __shared__ double sh_s[kBlock], sh_c[kBlock];
    double sum = 0.0, comp = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        double ss, e;
        two_sum(sum, parts[2 * b], ss, e);
        sum = ss;
        comp += parts[2 * b + 1] + e;
    }
    block_reduce_pair(sum, comp, sh_s, sh_c);
    if (threadIdx.x == 0) out[0] = (float)(sum + comp);
}

/* elec_normalize_source kernels.py:1289-1301, second loop: source *= I / max(total, 1e-12) */
__global__ void k_elec_scale(float *__restrict__ source, const float *__restrict__ total,
                             float target, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float t = total[0];
    float scale = target / (t > 1e-12f ? t : 1e-12f);
    source[c] *= scale;
}

/* elec_jacobi_step kernels.py:1315-1356 */
__global__ void k_elec_jacobi(const float *__restrict__ phi_in, float *__restrict__ phi_out,
                              const float *__restrict__ sigma, const float *__restrict__ source,
                              const uint8_t *__restrict__ flags, float dx, float omega,
                              int nz_solid, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    size_t c = d.idx(i, j, k);
    uint8_t f = flags[c];
    if (f == GAS || (k <= nz_solid && f == SOLID)) {
        phi_out[c] = 0.0f;
        return;
    }
    const float eps = 1e-12f;
    float sp = sigma_at(sigma, i + 1, j, k, d), sm = sigma_at(sigma, i - 1, j, k, d);
    float sn = sigma_at(sigma, i, j + 1, k, d), ss = sigma_at(sigma, i, j - 1, k, d);
    float st = sigma_at(sigma, i, j, k + 1, d), sb = sigma_at(sigma, i, j, k - 1, d);
    float denom = sp + sm + sn + ss + st + sb + eps;
    float pn = sp * phi_in[d.idx(imin(i + 1, d.nx - 1), j, k)] +
               sm * phi_in[d.idx(imax(i - 1, 0), j, k)] +
               sn * phi_in[d.idx(i, imin(j + 1, d.ny - 1), k)] +
               ss * phi_in[d.idx(i, imax(j - 1, 0), k)] +
               st * phi_in[d.idx(i, j, imin(k + 1, d.nz - 1))] +
               sb * phi_in[d.idx(i, j, imax(k - 1, 0))];
    float phi_new = (pn + source[c] * dx * dx) / denom;
    phi_out[c] = (1.0f - omega) * phi_in[c] + omega * phi_new;
}

/* elec_l1_diff kernels.py:1384-1396: compensated partial sums of |a - b| and |b| */
__global__ void __launch_bounds__(kBlock) k_elec_l1(const float *__restrict__ a,
                                                    const float *__restrict__ b,
                                                    double *__restrict__ parts, size_t N)
{
    __shared__ double sh_s[kBlock], sh_c[kBlock];
    double s0 = 0.0, c0 = 0.0, s1 = 0.0, c1 = 0.0;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N;
         c += (size_t)gridDim.x * blockDim.x) {
        float av = a[c], bv = b[c];
        double ss, e;
        two_sum(s0, (double)fabsf(av - bv), ss, e);
        s0 = ss;
        c0 += e;
        two_sum(s1, (double)fabsf(bv), ss, e);
        s1 = ss;
        c1 += e;
    }
    block_reduce_pair(s0, c0, sh_s, sh_c);
    if (threadIdx.x == 0) {
        parts[2 * blockIdx.x] = s0;
        parts[2 * blockIdx.x + 1] = c0;
    }
    __syncthreads();
    block_reduce_pair(s1, c1, sh_s, sh_c);
    if (threadIdx.x == 0) {
        parts[2 * (gridDim.x + blockIdx.x)] = s1;
        parts[2 * (gridDim.x + blockIdx.x) + 1] = c1;
    }
}

/* elec_compute_J kernels.py:1359-1381 */
__global__ void k_elec_J(float *__restrict__ Jx, float *__restrict__ Jy, float *__restrict__ Jz,
                         const float *__restrict__ phi, const float *__restrict__ sigma,
                         const uint8_t *__restrict__ flags, float dx, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    size_t c = d.idx(i, j, k);
    if (flags[c] == GAS) {
        Jx[c] = 0.0f;
        Jy[c] = 0.0f;
        Jz[c] = 0.0f;
        return;
    }
    float sig = sigma[c];
    float gx = (phi[d.idx(imin(i + 1, d.nx - 1), j, k)] - phi[d.idx(imax(i - 1, 0), j, k)]) /
               (2.0f * dx);
    float gy = (phi[d.idx(i, imin(j + 1, d.ny - 1), k)] - phi[d.idx(i, imax(j - 1, 0), k)]) /
               (2.0f * dx);
    float gz = (phi[d.idx(i, j, imin(k + 1, d.nz - 1))] - phi[d.idx(i, j, imax(k - 1, 0))]) /
               (2.0f * dx);
    Jx[c] = -sig * gx;
    Jy[c] = -sig * gy;
    Jz[c] = -sig * gz;
}

__device__ __forceinline__ int radial_bin(int i, int j, float arc_i, float arc_j, float dr,
                                          int n_bins)
{
    float di = (float)i - arc_i, dj = (float)j - arc_j;
    float r = sqrtf(di * di + dj * dj);
    return imin((int)(r / dr), n_bins - 1);
}

/* elec_bin_axial_current kernels.py:1399-1418: bins[k][b] = sum over the slab's cells of
 * radial bin b of Jz dx dx. Workgroup (k, g) sums bins g*kBinGroup.. of slab k: each thread
 * keeps compensated pairs for those bins over a strided part of the slab, then the workgroup
 * combines the threads in a fixed order. */
__global__ void __launch_bounds__(kBlock) k_elec_bins(const float *__restrict__ Jz,
                                                      float *__restrict__ bins, float arc_i,
                                                      float arc_j, float dx, float dr,
                                                      int n_bins, Dims d)
{
    __shared__ double sh_s[kBlock], sh_c[kBlock];
    int k = blockIdx.x;
    int b0 = blockIdx.y * kBinGroup;
    double s[kBinGroup], e[kBinGroup];
#pragma unroll
    for (int g = 0; g < kBinGroup; ++g) s[g] = e[g] = 0.0;
    int nij = d.nx * d.ny;
    for (int t = threadIdx.x; t < nij; t += blockDim.x) {
        int i = t % d.nx, j = t / d.nx;
        int b = radial_bin(i, j, arc_i, arc_j, dr, n_bins) - b0;
        if (b < 0 || b >= kBinGroup) continue;
        float v = Jz[d.idx(i, j, k)] * dx * dx;
#pragma unroll
        for (int g = 0; g < kBinGroup; ++g)
            if (g == b) {
                double ss, er;
                two_sum(s[g], (double)v, ss, er);
                s[g] = ss;
                e[g] += er;
            }
    }
#pragma unroll
    for (int g = 0; g < kBinGroup; ++g) {
        double a = s[g], c = e[g];
        block_reduce_pair(a, c, sh_s, sh_c);
        if (threadIdx.x == 0 && b0 + g < n_bins) bins[(size_t)k * n_bins + b0 + g] = (float)(a + c);
        __syncthreads();
    }
}

/* elec_prefix_bins kernels.py:1421-1432: running fp32 sum along the bins of each slab */
__global__ void k_elec_prefix(float *__restrict__ bins, int n_bins, int nz)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nz) return;
    float acc = 0.0f;
    for (int b = 0; b < n_bins; ++b) {
        acc += bins[(size_t)k * n_bins + b];
        bins[(size_t)k * n_bins + b] = acc;
    }
}

/* elec_B_axisymmetric kernels.py:1435-1477 */
__global__ void k_elec_B(float *__restrict__ Bx, float *__restrict__ By, float *__restrict__ Bz,
                         const float *__restrict__ bins, float arc_i, float arc_j, float dx,
                         float dr, float mu0, int n_bins, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    size_t c = d.idx(i, j, k);
    float di = (float)i - arc_i, dj = (float)j - arc_j;
    float r_cells = sqrtf(di * di + dj * dj);
    float bx = 0.0f, by = 0.0f;
    if (r_cells > 1e-3f) {
        int b = imin((int)(r_cells / dr), n_bins - 1);
        float I_enc = bins[(size_t)k * n_bins + b];
        float r_m = r_cells * dx;
        float B_theta = mu0 * I_enc / (2.0f * 3.14159265358979323846f * r_m);
        bx = -B_theta * dj / r_cells;
        by = B_theta * di / r_cells;
    }
    Bx[c] = bx;
    By[c] = by;
    Bz[c] = 0.0f;
}

/* apply_lorentz_JxB kernels.py:1480-1513 */
__global__ void k_lorentz_JxB(float *__restrict__ Fx, float *__restrict__ Fy,
                              float *__restrict__ Fz, const float *__restrict__ Jx,
                              const float *__restrict__ Jy, const float *__restrict__ Jz,
                              const float *__restrict__ Bx, const float *__restrict__ By,
                              const float *__restrict__ Bz, const float *__restrict__ f_l,
                              const uint8_t *__restrict__ flags, float dt, float dx, float rho,
                              size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float fl = f_l[c];
    if (flags[c] == GAS || fl < 0.05f) return;
    float jx = Jx[c], jy = Jy[c], jz = Jz[c], bx = Bx[c], by = By[c], bz = Bz[c];
    float fx = jy * bz - jz * by;
    float fy = jz * bx - jx * bz;
    float fz = jx * by - jy * bx;
    float flw = fl < 1.0f ? fl : 1.0f;
    float scale = flw * dt * dt / (rho * dx);
    Fx[c] += fx * scale;
    Fy[c] += fy * scale;
    Fz[c] += fz * scale;
}

}  // namespace waam

using namespace waam;

extern "C" {

int waam_elec_build_sigma(float *sigma, const float *f_l, const uint8_t *flags,
                          float sigma_liquid, float sigma_solid, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_elec_build_sigma: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_elec_build_sigma", {sigma, f_l, flags}));
    hipLaunchKernelGGL(k_elec_build_sigma, dim3(grid_for((size_t)n)), dim3(kBlock), 0,
                       as_stream(stream), sigma, f_l, flags, sigma_liquid, sigma_solid, (size_t)n);
    return check_launch("waam_elec_build_sigma");
}

int waam_elec_clear_source(float *source, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_elec_clear_source: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_elec_clear_source", {source}));
    hipLaunchKernelGGL(k_fill, dim3(grid_for((size_t)n)), dim3(kBlock), 0, as_stream(stream),
                       source, 0.0f, (size_t)n);
    return check_launch("waam_elec_clear_source");
}

int waam_elec_inject_arc_source(float *source, float arc_i, float arc_j, float arc_k,
                                float sigma_cells, float current_A, float dx, int nx, int ny,
                                int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_elec_inject_arc_source", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_elec_inject_arc_source", {source}));
    Dims d{nx, ny, nz, layout};
    ElecSource e;
    e.arc_i = arc_i;
    e.arc_j = arc_j;
    e.arc_k = arc_k;
    e.inv2s2 = 1.0f / (2.0f * sigma_cells * sigma_cells + 1e-6f);
    e.current = current_A;
    e.cell_vol = dx * dx * dx;
    int km = (int)(arc_k + sigma_cells * 4.0f) + 1;
    e.k_max = km < nz ? km : nz;
    int kn = (int)(arc_k - sigma_cells);
    e.k_min = kn > 0 ? kn : 0;
    hipLaunchKernelGGL(k_elec_source, grid3(d), block3(), 0, as_stream(stream), source, e, d);
    return check_launch("waam_elec_inject_arc_source");
}

int waam_elec_normalize_source(float *source, float target_current_A, float dx, int64_t n,
                               void *scratch, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_elec_normalize_source: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_elec_normalize_source", {source, scratch}));
    hipStream_t s = as_stream(stream);
    double *parts = reinterpret_cast<double *>((char *)scratch + kScratchPartials);
    float *total = reinterpret_cast<float *>((char *)scratch + kScratchTelemetry) + 8;
    hipLaunchKernelGGL(k_elec_source_sum, dim3(kLorentzSumBlocks), dim3(kBlock), 0, s, source,
                       dx * dx * dx, parts, (size_t)n);
    hipLaunchKernelGGL(k_pairs_total, dim3(1), dim3(kBlock), 0, s, parts, kLorentzSumBlocks,
                       total);
    hipLaunchKernelGGL(k_elec_scale, dim3(grid_for((size_t)n)), dim3(kBlock), 0, s, source, total,
                       target_current_A, (size_t)n);
    return check_launch("waam_elec_normalize_source");
}

int waam_elec_init_ground(float *phi, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_elec_init_ground: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_elec_init_ground", {phi}));
    hipLaunchKernelGGL(k_fill, dim3(grid_for((size_t)n)), dim3(kBlock), 0, as_stream(stream), phi,
                       0.0f, (size_t)n);
    return check_launch("waam_elec_init_ground");
}

int waam_elec_jacobi_step(const float *phi_in, float *phi_out, const float *sigma,
                          const float *source, const uint8_t *flags, float dx, float omega_j,
                          int nz_solid, int nx, int ny, int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_elec_jacobi_step", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_elec_jacobi_step", {phi_in, phi_out, sigma, source, flags}));
    if (phi_in == phi_out)
        return set_error(WAAM_ERR_INVALID, "waam_elec_jacobi_step: phi_in aliases phi_out");
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_elec_jacobi, grid3(d), block3(), 0, as_stream(stream), phi_in, phi_out,
                       sigma, source, flags, dx, omega_j, nz_solid, d);
    return check_launch("waam_elec_jacobi_step");
}

int waam_elec_l1_diff(const float *a, const float *b, float *out, int64_t n, void *scratch,
                      void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_elec_l1_diff: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_elec_l1_diff", {a, b, out, scratch}));
    hipStream_t s = as_stream(stream);
    double *parts = reinterpret_cast<double *>((char *)scratch + kScratchTelemetryParts);
    const int nb = kLorentzSumBlocks;
    hipLaunchKernelGGL(k_elec_l1, dim3(nb), dim3(kBlock), 0, s, a, b, parts, (size_t)n);
    hipLaunchKernelGGL(k_pairs_total, dim3(1), dim3(kBlock), 0, s, parts, nb, out);
    hipLaunchKernelGGL(k_pairs_total, dim3(1), dim3(kBlock), 0, s, parts + 2 * nb, nb, out + 1);
    return check_launch("waam_elec_l1_diff");
}

int waam_elec_compute_J(float *Jx, float *Jy, float *Jz, const float *phi, const float *sigma,
                        const uint8_t *flags, float dx, int nx, int ny, int nz, int layout,
                        void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_elec_compute_J", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_elec_compute_J", {Jx, Jy, Jz, phi, sigma, flags}));
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_elec_J, grid3(d), block3(), 0, as_stream(stream), Jx, Jy, Jz, phi, sigma,
                       flags, dx, d);
    return check_launch("waam_elec_compute_J");
}

int waam_elec_bin_axial_current(const float *Jz, float *bins, float arc_i, float arc_j, float dx,
                                float bin_dr_cells, int n_bins, int nx, int ny, int nz,
                                int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_elec_bin_axial_current", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_elec_bin_axial_current", {Jz, bins}));
    if (n_bins <= 0 || n_bins > 4096)
        return set_error(WAAM_ERR_INVALID, "waam_elec_bin_axial_current: n_bins %d", n_bins);
    if (!(bin_dr_cells > 0.0f))
        return set_error(WAAM_ERR_INVALID, "waam_elec_bin_axial_current: bin width <= 0");
    Dims d{nx, ny, nz, layout};
    dim3 g((unsigned)nz, (unsigned)((n_bins + kBinGroup - 1) / kBinGroup));
    hipLaunchKernelGGL(k_elec_bins, g, dim3(kBlock), 0, as_stream(stream), Jz, bins, arc_i, arc_j,
                       dx, bin_dr_cells, n_bins, d);
    return check_launch("waam_elec_bin_axial_current");
}

int waam_elec_prefix_bins(float *bins, int n_bins, int nz, void *stream)
{
    if (n_bins <= 0 || nz <= 0) return set_error(WAAM_ERR_INVALID, "waam_elec_prefix_bins: size");
    WAAM_RETURN_IF(check_ptrs("waam_elec_prefix_bins", {bins}));
    hipLaunchKernelGGL(k_elec_prefix, dim3(grid_for((size_t)nz, 64)), dim3(64), 0,
                       as_stream(stream), bins, n_bins, nz);
    return check_launch("waam_elec_prefix_bins");
}

int waam_elec_B_axisymmetric(float *Bx, float *By, float *Bz, const float *bins, float arc_i,
                             float arc_j, float dx, float bin_dr_cells, float mu0, int n_bins,
                             int nx, int ny, int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_elec_B_axisymmetric", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_elec_B_axisymmetric", {Bx, By, Bz, bins}));
    if (n_bins <= 0 || !(bin_dr_cells > 0.0f))
        return set_error(WAAM_ERR_INVALID, "waam_elec_B_axisymmetric: bins");
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_elec_B, grid3(d), block3(), 0, as_stream(stream), Bx, By, Bz, bins, arc_i,
                       arc_j, dx, bin_dr_cells, mu0, n_bins, d);
    return check_launch("waam_elec_B_axisymmetric");
}

int waam_apply_lorentz_JxB(float *Fx, float *Fy, float *Fz, const float *Jx, const float *Jy,
                           const float *Jz, const float *Bx, const float *By, const float *Bz,
                           const float *f_l, const uint8_t *flags, float dt, float dx,
                           float rho_ref, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_apply_lorentz_JxB: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_apply_lorentz_JxB", {Fx, Fy, Fz, Jx, Jy, Jz, Bx, By, Bz, f_l,
                                                         flags}));
    hipLaunchKernelGGL(k_lorentz_JxB, dim3(grid_for((size_t)n)), dim3(kBlock), 0,
                       as_stream(stream), Fx, Fy, Fz, Jx, Jy, Jz, Bx, By, Bz, f_l, flags, dt, dx,
                       rho_ref, (size_t)n);
    return check_launch("waam_apply_lorentz_JxB");
}

}  // extern "C"
