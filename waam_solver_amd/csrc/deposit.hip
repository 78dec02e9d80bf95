// deposit.hip — wire-droplet deposition, porosity tracers and the pool telemetry reduction.
//
// Deposition (reference kernels.py:245-400, host retry loop solvers/coupled_step.py:126-151).
// The reference kernel sweeps a sub-box three times, converting GAS cells to FLUID under a
// reserve-then-commit volume budget; whether a cell qualifies depends on whether a neighbour
// was converted earlier in the same sweep. Its pinned semantics here are those of the sweep in
// i, j, k order. Under that order a cell's qualification depends only on its three
// lexicographic predecessors (i-1,j,k), (i,j-1,k), (i,j,k-1), which lie on the previous
// anti-diagonal of the box; and the budget admits exactly the first n qualifying cells, n being
// the number of fp32 increments of the accumulator that stay below the target. One workgroup
// therefore (1) marks qualifying cells diagonal by diagonal, assuming an unlimited budget,
// (2) ranks them in i, j, k order with a block scan, (3) converts the first n. All retries of
// the host loop run inside the same launch.
#include "waam_common.h"

namespace waam {

constexpr int kDepThreads = 1024;

struct DepositArgs {
    float arc_i, arc_j, arc_k, droplet_radius, T_drop, cp_rho, L_rho, rho0, cell_vol;
    float target_vol_f;   /* fp32 budget target passed to each sweep */
    double target_vol;    /* host value for the 0.98 acceptance test */
    double r_try0;        /* first footprint radius (host double) */
    int max_tries;        /* 1: a single reference kernel call; 10: the host retry loop */
    Dims d;
    const float *arc_k_dev; /* non-NULL: arc_k is read here (resolved on the device) */
};

__device__ __forceinline__ bool is_metal(const uint8_t *flags, const float *phi, const float *f_l,
                                         size_t c)
{
    uint8_t f = flags[c];
    return f == FLUID || f == SOLID || phi[c] > 0.35f || f_l[c] > 0.2f;
}

__global__ void __launch_bounds__(kDepThreads)
    k_deposit(float *__restrict__ f_src, uint8_t *__restrict__ flags, float *__restrict__ f_l,
              float *__restrict__ phi, float *__restrict__ H, float *__restrict__ T,
              float *__restrict__ rho, float *__restrict__ vol_acc, float *__restrict__ real_acc,
              uint8_t *__restrict__ mark, float *__restrict__ result, DepositArgs a)
{
    __shared__ int s_cnt[kDepThreads];
    __shared__ int s_allowed;
    __shared__ float s_acc, s_real;
    __shared__ int s_done;
    __shared__ double s_rtry, s_placed;
    const Dims d = a.d;
    const size_t N = d.n();
    const int tid = threadIdx.x;
    const float arc_k = a.arc_k_dev ? *a.arc_k_dev : a.arc_k;
    int tries = 0;
    if (tid == 0) {
        s_rtry = a.r_try0;
        s_placed = 0.0;
    }
    __syncthreads();
    for (int attempt = 0; attempt < a.max_tries; ++attempt) {
        const double placed = s_placed;
        if (tid == 0) {
            if (a.max_tries > 1) {
                vol_acc[0] = (float)placed;
                real_acc[0] = (float)placed;
            }
            s_acc = vol_acc[0];
            s_real = real_acc[0];
        }
        __syncthreads();
        const float footprint_r = (float)s_rtry;
        const float dr = a.droplet_radius;
        const float search_r = fmx(footprint_r, dr);
        const int i_min = imax(0, (int)(a.arc_i - search_r));
        const int i_max = imin(d.nx, (int)(a.arc_i + search_r) + 1);
        const int j_min = imax(0, (int)(a.arc_j - search_r));
        const int j_max = imin(d.ny, (int)(a.arc_j + search_r) + 1);
        const int k_lo = imax(0, (int)(arc_k - 1.0f));
        const float k_span = fmx(dr + 3.0f, 4.0f);
        const int k_hi = imin(d.nz, (int)(arc_k + k_span) + 1);
        const float drop_cz = arc_k + dr + 0.5f;
        const float r_drop2 = dr * dr;
        const float foot2 = search_r * search_r;
        const int bi = i_max - i_min, bj = j_max - j_min, bk = k_hi - k_lo;
        const int B = (bi > 0 && bj > 0 && bk > 0) ? bi * bj * bk : 0;

        for (int pass = 1; pass <= 2 && B > 0; ++pass) {
            /* (1) qualification under an unlimited budget, one anti-diagonal at a time */
            for (int t = tid; t < B; t += kDepThreads) mark[t] = 0;
            __syncthreads();
            const int ndiag = bi + bj + bk - 2;
            for (int dg = 0; dg < ndiag; ++dg) {
                for (int t = tid; t < B; t += kDepThreads) {
                    int lk = t % bk, lj = (t / bk) % bj, li = t / (bk * bj);
                    if (li + lj + lk != dg) continue;
                    int i = i_min + li, j = j_min + lj, k = k_lo + lk;
                    size_t c = d.idx(i, j, k);
                    if (flags[c] != GAS) continue;
                    float di = (float)i - a.arc_i, dj = (float)j - a.arc_j;
                    float rij2 = di * di + dj * dj;
                    if (pass == 1) {
                        float dk = (float)k - drop_cz;
                        float r2 = di * di + dj * dj + dk * dk;
                        if (rij2 > foot2 && r2 > r_drop2) continue;
                    } else {
                        if (rij2 > foot2 * 1.44f) continue;
                    }
                    /* metal neighbour: converted predecessor in this sweep, or metal now */
                    bool metal = false;
                    const int nb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0},
                                          {0, 1, 0},  {0, 0, -1}, {0, 0, 1}};
                    for (int q = 0; q < 6 && !metal; ++q) {
                        int ii = i + nb[q][0], jj = j + nb[q][1], kk = k + nb[q][2];
                        if (ii < 0 || jj < 0 || kk < 0 || ii >= d.nx || jj >= d.ny || kk >= d.nz)
                            continue;
                        int mi = ii - i_min, mj = jj - j_min, mk = kk - k_lo;
                        if (mi >= 0 && mi < bi && mj >= 0 && mj < bj && mk >= 0 && mk < bk &&
                            mark[(mi * bj + mj) * bk + mk])
                            metal = true;
                        else if (is_metal(flags, phi, f_l, d.idx(ii, jj, kk)))
                            metal = true;
                    }
                    if (metal) mark[t] = 1;
                }
                __threadfence_block();
                __syncthreads();
            }
            /* (2) rank in i, j, k order: contiguous chunks per thread, block scan of counts */
            const int chunk = (B + kDepThreads - 1) / kDepThreads;
            const int t0 = tid * chunk, t1 = imin(B, t0 + chunk);
            int cnt = 0;
            for (int t = t0; t < t1; ++t) cnt += mark[t];
            s_cnt[tid] = cnt;
            __syncthreads();
            for (int off = 1; off < kDepThreads; off <<= 1) {
                int v = tid >= off ? s_cnt[tid - off] : 0;
                __syncthreads();
                s_cnt[tid] += v;
                __syncthreads();
            }
            if (tid == kDepThreads - 1) {
                /* (3) budget: a cell converts while the fp32 accumulator is below target */
                int total = s_cnt[tid];
                float acc = s_acc, real = s_real;
                int n = 0;
                while (n < total && acc < a.target_vol_f) {
                    acc = acc + a.cell_vol;
                    real = real + a.cell_vol;
                    ++n;
                }
                s_allowed = n;
                s_acc = acc;
                s_real = real;
            }
            __syncthreads();
            int rank = s_cnt[tid] - cnt;
            const int allowed = s_allowed;
            for (int t = t0; t < t1; ++t) {
                if (!mark[t]) continue;
                if (rank < allowed) {
                    int lk = t % bk, lj = (t / bk) % bj, li = t / (bk * bj);
                    size_t c = d.idx(i_min + li, j_min + lj, k_lo + lk);
                    flags[c] = FLUID;
                    f_l[c] = 1.0f;
                    phi[c] = 1.0f;
                    T[c] = a.T_drop;
                    H[c] = a.cp_rho * a.T_drop + a.L_rho;
                    rho[c] = a.rho0;
#pragma unroll
                    for (int q = 0; q < Q; ++q) f_src[(size_t)q * N + c] = wq(q) * a.rho0;
                }
                ++rank;
            }
            __threadfence_block();
            __syncthreads();
        }
        /* pass 3: thermal top-up of liquid footprint cells (order independent) */
        for (int t = tid; t < B; t += kDepThreads) {
            int lk = t % bk, lj = (t / bk) % bj, li = t / (bk * bj);
            int i = i_min + li, j = j_min + lj;
            size_t c = d.idx(i, j, k_lo + lk);
            if (flags[c] != FLUID || f_l[c] < 0.45f) continue;
            float di = (float)i - a.arc_i, dj = (float)j - a.arc_j;
            if (di * di + dj * dj > foot2 * 1.44f) continue;
            T[c] = a.T_drop;
            H[c] = a.cp_rho * a.T_drop + a.L_rho;
            rho[c] = a.rho0;
#pragma unroll
            for (int q = 0; q < Q; ++q) f_src[(size_t)q * N + c] = wq(q) * a.rho0;
        }
        __threadfence_block();
        __syncthreads();
        tries = attempt + 1;
        if (tid == 0) {
            vol_acc[0] = s_acc;
            real_acc[0] = s_real;
            s_placed = (double)s_real;
            s_done = (a.max_tries == 1 || s_placed >= a.target_vol * 0.98) ? 1 : 0;
            if (!s_done) s_rtry = fmin(s_rtry + 1.5, 16.0);
        }
        __syncthreads();
        if (s_done) break;
    }
    if (tid == 0 && result) {
        result[0] = (float)s_placed;
        result[1] = (float)tries;
    }
}

/* inject_tracers — reference kernels.py:3182-3204; ti.random replaced by a counter hash */
__device__ __forceinline__ uint32_t hash_u32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ void k_inject_tracers(float *__restrict__ pos, int32_t *__restrict__ active,
                                 int32_t *__restrict__ head, int max_tracers, float ax, float ay,
                                 float az, float sigma_m, int spawn, uint32_t seed,
                                 const float *__restrict__ arc_k_dev, double dx)
{
    if (threadIdx.x != 0) return;
    /* arc_k resolved on the device: the height (arc_k + 1) * dx as the host forms it, in
     * binary64, rounded once */
    if (arc_k_dev) az = (float)(((double)*arc_k_dev + 1.0) * dx);
    for (int s = 0; s < spawn; ++s) {
        int32_t h = head[0];
        head[0] = h + 1;
        int slot = (int)((uint32_t)h % (uint32_t)max_tracers);
        active[slot] = 1;
        uint32_t base = hash_u32(seed ^ (uint32_t)h * 2654435761u);
        float r1 = (float)(hash_u32(base + 1u) >> 8) * (1.0f / 16777216.0f);
        float r2 = (float)(hash_u32(base + 2u) >> 8) * (1.0f / 16777216.0f);
        pos[3 * slot + 0] = ax + (r1 * 2.0f - 1.0f) * sigma_m;
        pos[3 * slot + 1] = ay + (r2 * 2.0f - 1.0f) * sigma_m;
        pos[3 * slot + 2] = az;
    }
}

/* advect_tracers — reference kernels.py:3206-3255 */
__global__ void k_advect_tracers(float *__restrict__ pos, int32_t *__restrict__ active,
                                 const float *__restrict__ ux, const float *__restrict__ uy,
                                 const float *__restrict__ uz, const float *__restrict__ f_l,
                                 const uint8_t *__restrict__ flags, float dx, int max_tracers,
                                 Dims d)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= max_tracers || active[p] != 1) return;
    float px = pos[3 * p] / dx, py = pos[3 * p + 1] / dx, pz = pos[3 * p + 2] / dx;
    int i = (int)px, j = (int)py, k = (int)pz;
    if (i < 0 || i >= d.nx || j < 0 || j >= d.ny || k < 0 || k >= d.nz) {
        active[p] = 0;
        return;
    }
    size_t c = d.idx(i, j, k);
    uint8_t f = flags[c];
    if (f == GAS) {
        active[p] = 0;
        return;
    }
    if (f_l[c] < 0.05f || f == SOLID) {
        active[p] = 2;
        return;
    }
    pos[3 * p] += ux[c] * dx;
    pos[3 * p + 1] += uy[c] * dx;
    pos[3 * p + 2] += uz[c] * dx;
}

/* telemetry_pool_reduce — reference kernels.py:3424-3468. Max/min exact; the liquid count and
 * the i sum are exact integers (the reference's float atomics become order independent). */
struct Telem {
    double n, Tp, cool, u2, isum, imin, imax, Tg;
};

__device__ __forceinline__ void telem_merge(Telem &a, const Telem &b)
{
    a.n += b.n;
    a.Tp = fmax(a.Tp, b.Tp);
    a.cool = fmax(a.cool, b.cool);
    a.u2 = fmax(a.u2, b.u2);
    a.isum += b.isum;
    a.imin = fmin(a.imin, b.imin);
    a.imax = fmax(a.imax, b.imax);
    a.Tg = fmax(a.Tg, b.Tg);
}

__device__ void telem_block_reduce(Telem &v, Telem *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) telem_merge(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    v = sh[0];
}

__global__ void __launch_bounds__(kBlock) k_telemetry(const float *__restrict__ T,
                                                      const float *__restrict__ f_l,
                                                      const float *__restrict__ phi,
                                                      const float *__restrict__ dT_dt,
                                                      const float *__restrict__ ux,
                                                      const float *__restrict__ uy,
                                                      const float *__restrict__ uz, Dims d,
                                                      Telem *__restrict__ parts)
{
    __shared__ Telem sh[kBlock];
    Telem v{0.0, 0.0, 0.0, 0.0, 0.0, 1.0e9, -1.0e9, 0.0};
    size_t N = d.n();
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N;
         c += (size_t)gridDim.x * blockDim.x) {
        float Tc = T[c];
        v.Tg = fmax(v.Tg, (double)Tc);
        if (phi[c] > 0.05f) v.cool = fmax(v.cool, (double)(-dT_dt[c]));
        if (f_l[c] > 0.5f) {
            int ci, cj, ck;
            d.ijk(c, ci, cj, ck);
            double i = (double)ci;
            v.n += 1.0;
            v.Tp = fmax(v.Tp, (double)Tc);
            float u2 = ux[c] * ux[c] + uy[c] * uy[c] + uz[c] * uz[c];
            v.u2 = fmax(v.u2, (double)u2);
            v.isum += i;
            v.imin = fmin(v.imin, i);
            v.imax = fmax(v.imax, i);
        }
    }
    telem_block_reduce(v, sh);
    if (threadIdx.x == 0) parts[blockIdx.x] = v;
}

__global__ void __launch_bounds__(kBlock) k_telemetry_final(const Telem *__restrict__ parts,
                                                            int nparts, double *__restrict__ out)
{
    __shared__ Telem sh[kBlock];
    Telem v{0.0, 0.0, 0.0, 0.0, 0.0, 1.0e9, -1.0e9, 0.0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) telem_merge(v, parts[b]);
    telem_block_reduce(v, sh);
    if (threadIdx.x == 0) {
        out[0] = v.n;
        out[1] = v.Tp;
        out[2] = v.cool;
        out[3] = v.u2;
        out[4] = v.isum;
        out[5] = v.imin;
        out[6] = v.imax;
        out[7] = v.Tg;
    }
}

}  // namespace waam

using namespace waam;

static int deposit_launch(const char *name, float *f_src, uint8_t *flags, float *f_l, float *phi,
                          float *H, float *T, float *rho, float *vol_acc, float *real_acc,
                          void *scratch, float *result, const DepositArgs &a, void *stream)
{
    WAAM_RETURN_IF(check_dims(name, a.d.nx, a.d.ny, a.d.nz, a.d.layout));
    WAAM_RETURN_IF(check_ptrs(name, {f_src, flags, f_l, phi, H, T, rho, vol_acc, real_acc, scratch}));
    double sr = a.r_try0 > (double)a.droplet_radius ? a.r_try0 : (double)a.droplet_radius;
    if (a.max_tries > 1 && sr < 16.0) sr = 16.0;
    double span_i = 2.0 * sr + 2.0, span_k = (a.droplet_radius + 3.0 > 4.0 ? a.droplet_radius + 3.0 : 4.0) + 3.0;
    if (span_i * span_i * span_k > (double)kDepMaxCells)
        return set_error(WAAM_ERR_INVALID, "%s: deposition box of radius %.1f exceeds the workspace",
                         name, sr);
    uint8_t *mark = reinterpret_cast<uint8_t *>((char *)scratch + kDepScratchOffset);
    hipLaunchKernelGGL(k_deposit, dim3(1), dim3(kDepThreads), 0, as_stream(stream), f_src, flags,
                       f_l, phi, H, T, rho, vol_acc, real_acc, mark, result, a);
    return check_launch(name);
}

namespace waam {

int launch_deposit_dev(float *f_src, uint8_t *flags, float *f_l, float *phi, float *H, float *T,
                       float *rho, float *vol_acc, float *real_acc, void *scratch, float *result,
                       float arc_i, float arc_j, const float *arc_k_dev, double footprint_r,
                       float droplet_radius, double drop_vol, float T_drop, float cp_rho,
                       float L_rho, float rho0, float cell_vol, Dims d, hipStream_t st)
{
    DepositArgs a{};
    a.arc_i = arc_i;
    a.arc_j = arc_j;
    a.arc_k = 0.0f;
    a.arc_k_dev = arc_k_dev;
    a.droplet_radius = droplet_radius;
    a.T_drop = T_drop;
    a.cp_rho = cp_rho;
    a.L_rho = L_rho;
    a.rho0 = rho0;
    a.cell_vol = cell_vol;
    a.target_vol_f = (float)drop_vol;
    a.target_vol = drop_vol;
    a.r_try0 = footprint_r;
    a.max_tries = 10;
    a.d = d;
    return deposit_launch("waam_step_fused_droplet", f_src, flags, f_l, phi, H, T, rho, vol_acc,
                          real_acc, scratch, result, a, st);
}

int launch_tracers_dev(float *pos, int32_t *active, int32_t *head, int max_tracers, float ax,
                       float ay, const float *arc_k_dev, double dx, float sigma_m, int spawn,
                       uint32_t seed, hipStream_t st)
{
    hipLaunchKernelGGL(k_inject_tracers, dim3(1), dim3(64), 0, st, pos, active, head,
                       max_tracers, ax, ay, 0.0f, sigma_m, spawn, seed, arc_k_dev, dx);
    return check_launch("waam_step_fused_droplet (tracers)");
}

}  // namespace waam

extern "C" {

int waam_feed_wire_surface(float *f_src, uint8_t *flags, float *f_l, float *phi, float *H,
                           float *T, float *rho, float arc_i, float arc_j, float arc_k,
                           float footprint_r, float droplet_radius, float target_vol,
                           float T_drop, float cp_rho, float L_rho, float rho0, float *vol_acc,
                           float *real_acc, float cell_vol, int nx, int ny, int nz, int layout,
                           void *scratch, void *stream)
{
    DepositArgs a{};
    a.arc_i = arc_i;
    a.arc_j = arc_j;
    a.arc_k = arc_k;
    a.droplet_radius = droplet_radius;
    a.T_drop = T_drop;
    a.cp_rho = cp_rho;
    a.L_rho = L_rho;
    a.rho0 = rho0;
    a.cell_vol = cell_vol;
    a.target_vol_f = target_vol;
    a.target_vol = (double)target_vol;
    a.r_try0 = (double)footprint_r;
    a.max_tries = 1;
    a.d = Dims{nx, ny, nz, layout};
    return deposit_launch("waam_feed_wire_surface", f_src, flags, f_l, phi, H, T, rho, vol_acc,
                          real_acc, scratch, nullptr, a, stream);
}

int waam_deposit_droplet(float *f_src, uint8_t *flags, float *f_l, float *phi, float *H,
                         float *T, float *rho, float arc_i, float arc_j, float arc_k,
                         double footprint_r, float droplet_radius, double drop_vol,
                         float T_drop, float cp_rho, float L_rho, float rho0, float *vol_acc,
                         float *real_acc, float cell_vol, float *result, int nx, int ny, int nz, int layout,
                         void *scratch, void *stream)
{
    DepositArgs a{};
    a.arc_i = arc_i;
    a.arc_j = arc_j;
    a.arc_k = arc_k;
    a.droplet_radius = droplet_radius;
    a.T_drop = T_drop;
    a.cp_rho = cp_rho;
    a.L_rho = L_rho;
    a.rho0 = rho0;
    a.cell_vol = cell_vol;
    a.target_vol_f = (float)drop_vol;
    a.target_vol = drop_vol;
    a.r_try0 = footprint_r;
    a.max_tries = 10;
    a.d = Dims{nx, ny, nz, layout};
    return deposit_launch("waam_deposit_droplet", f_src, flags, f_l, phi, H, T, rho, vol_acc,
                          real_acc, scratch, result, a, stream);
}

int waam_inject_tracers(float *pos, int32_t *active, int32_t *head, int max_tracers,
                        float arc_x, float arc_y, float arc_z, float sigma_m, int spawn_count,
                        uint32_t seed, void *stream)
{
    if (max_tracers <= 0) return set_error(WAAM_ERR_INVALID, "waam_inject_tracers: max_tracers <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_inject_tracers", {pos, active, head}));
    hipLaunchKernelGGL(k_inject_tracers, dim3(1), dim3(64), 0, as_stream(stream), pos, active,
                       head, max_tracers, arc_x, arc_y, arc_z, sigma_m, spawn_count, seed,
                       (const float *)nullptr, 0.0);
    return check_launch("waam_inject_tracers");
}

int waam_advect_tracers(float *pos, int32_t *active, const float *ux, const float *uy,
                        const float *uz, const float *f_l, const uint8_t *flags, float dx,
                        float dt, int max_tracers, int nx, int ny, int nz, int layout, void *stream)
{
    (void)dt;
    if (max_tracers <= 0) return WAAM_OK;
    WAAM_RETURN_IF(check_dims("waam_advect_tracers", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_advect_tracers", {pos, active, ux, uy, uz, f_l, flags}));
    hipLaunchKernelGGL(k_advect_tracers, dim3(grid_for((size_t)max_tracers)), dim3(kBlock), 0,
                       as_stream(stream), pos, active, ux, uy, uz, f_l, flags, dx, max_tracers,
                       Dims{nx, ny, nz, layout});
    return check_launch("waam_advect_tracers");
}

int waam_telemetry_pool_reduce(const float *T, const float *f_l, const float *phi,
                               const float *dT_dt, const float *ux, const float *uy,
                               const float *uz, double *out, int nx, int ny, int nz, int layout,
                               void *scratch, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_telemetry_pool_reduce", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_telemetry_pool_reduce", {T, f_l, phi, dT_dt, ux, uy, uz, out, scratch}));
    Telem *parts = reinterpret_cast<Telem *>((char *)scratch + kScratchTelemetryParts);
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_telemetry, dim3(kReduceBlocks), dim3(kBlock), 0, s, T, f_l, phi, dT_dt, ux,
                       uy, uz, Dims{nx, ny, nz, layout}, parts);
    hipLaunchKernelGGL(k_telemetry_final, dim3(1), dim3(kBlock), 0, s, parts, kReduceBlocks, out);
    return check_launch("waam_telemetry_pool_reduce");
}

}  // extern "C"
