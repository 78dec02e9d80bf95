// fused.hip — the coupled step in eight launches (waam_step_fused).
//
// Launch sequence for one step (reference solvers/coupled_step.py:73-464):
//   1 k_step_arc_partials   torch-column surface height (arc_k) + arc weight partial sums
//   2 k_step_arc_total      fixed-order weight total, energy per unit weight
//   3 k_step_thermal        arc + enthalpy ceiling applied on the fly, advection-diffusion,
//                           boundary losses, floor, ceiling, phase map, T_max, cooling rate
//   4 k_step_vof_advect     donor-cell phi advection + re-initialisation -> phi_tmp
//   5 k_step_vof_flags      contact-angle BC, flag classification, phi/f_l consistency, time
//                           above the HAZ thresholds
//   6 k_step_collide        CSF + Marangoni + arc pressure + gravity + buoyancy, force clamp,
//                           MRT/SRT collision, Mach clamp, and the previous step's pull stream
//   7 k_advect_tracers      porosity tracers
//   8 k_step_freeze         remelt, trailing-pool and cooled-metal solidification
// Every cell value is produced by the same fp32 expression sequence as the one-kernel-per-
// reference-kernel path, so both paths give identical bits.
//
// Deferred stream. The reference streams populations after each collide (f_dst -> f_src) and
// then swaps, so the next collide overwrites the streamed value on every FLUID/IFACE cell. Here
// the stream of step n is folded into kernel 6 of step n+1: a cell that step n+1 collides
// skips the pull (its value would be overwritten), every other cell that was not SOLID at step
// n pulls from its neighbours with step n's flags (kept in sflags). Populations are read once
// and written once per active cell per step instead of twice.
#include <algorithm>
#include <cmath>
#include <vector>

#include "arc_cell.h"
#include "forces_cell.h"
#include "lbm_cell.h"
#include "thermal_cell.h"
#include "vof_cell.h"

namespace waam {

struct StepConsts {
    waam_step_params p;
    ArcParams arc;
    PhaseConsts pc;
    LossConsts lc;
    ArcPressure ap;
    float cap_l, cap_s, h_min, inv_dt;
    int box_i0, box_i1, box_j0, box_j1, box_k0, box_k1;  // cells outside get exactly zero arc heat
    Dims d;
};

__device__ __forceinline__ float step_arc_k(const waam_step_params &p, const float *tot)
{
    return p.use_arc_k_override ? p.arc_k_override : tot[2];
}

/* 1 — surface height in the torch column (reference kernels.py:625-641 and
 *     coupled_step.py:34-59) and per-block compensated arc weight sums. */
__global__ void __launch_bounds__(kBlock) k_step_arc_partials(waam_step_fields f, StepConsts s,
                                                               int do_sums)
{
    __shared__ double sh_s[kBlock], sh_c[kBlock];
    __shared__ int sh_k[kBlock];
    __shared__ float s_arck;
    const Dims &d = s.d;
    const waam_step_params &p = s.p;
    int best = -1;
    for (int k = threadIdx.x; k < d.nz; k += blockDim.x) {
        size_t c = d.idx(p.col_i, p.col_j, k);
        if (f.flags_in[c] != GAS && f.phi[c] > 0.05f) best = k > best ? k : best;
    }
    sh_k[threadIdx.x] = best;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh_k[threadIdx.x] = imax(sh_k[threadIdx.x], sh_k[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int kf = imax(1, p.nz_solid - 1);
        float ks = (float)(sh_k[0] >= 0 ? sh_k[0] : kf);
        if (ks < 1.0f) ks = (float)imax(1, p.nz_solid - 1);
        if (p.use_k_robot) ks = fmx(ks, p.k_robot);
        s_arck = ks;
        if (blockIdx.x == 0) {
            float *tot = reinterpret_cast<float *>((char *)f.scratch + kScratchTelemetry);
            tot[2] = ks;
        }
    }
    __syncthreads();
    if (!do_sums) return;
    ArcParams ap = s.arc;
    ap.arc_k = s_arck;
    double sum = 0.0, comp = 0.0;
    size_t N = d.n();
    for (size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x; cell < N;
         cell += (size_t)gridDim.x * blockDim.x) {
        if (f.flags_in[cell] == GAS) continue;
        int k = (int)(cell % (size_t)d.nz);
        size_t r = cell / (size_t)d.nz;
        int j = (int)(r % (size_t)d.ny);
        int i = (int)(r / (size_t)d.ny);
        float wg, wdep;
        if (!arc_cell_weights(ap, i, j, k, f.fl_in[cell], wg, wdep)) continue;
        double ss, e;
        two_sum(sum, (double)(wg * wdep), ss, e);
        sum = ss;
        comp += e;
    }
    block_reduce_pair(sum, comp, sh_s, sh_c);
    if (threadIdx.x == 0) {
        double *parts = reinterpret_cast<double *>((char *)f.scratch + kScratchPartials);
        parts[2 * blockIdx.x] = sum;
        parts[2 * blockIdx.x + 1] = comp;
    }
}

/* 2 — fixed-order total of the block sums (reference kernels.py:493-496) */
__global__ void __launch_bounds__(kBlock) k_step_arc_total(waam_step_fields f, StepConsts s)
{
    __shared__ double sh_s[kBlock], sh_c[kBlock];
    const double *parts = reinterpret_cast<const double *>((char *)f.scratch + kScratchPartials);
    double sum = 0.0, comp = 0.0;
    for (int b = threadIdx.x; b < kReduceBlocks; b += blockDim.x) {
        double ss, e;
        two_sum(sum, parts[2 * b], ss, e);
        sum = ss;
        comp += parts[2 * b + 1] + e;
    }
    block_reduce_pair(sum, comp, sh_s, sh_c);
    if (threadIdx.x == 0) {
        float *tot = reinterpret_cast<float *>((char *)f.scratch + kScratchTelemetry);
        float w_total = (float)(sum + comp);
        float epw = 0.0f;
        if (w_total > 1e-9f) epw = s.p.Q_w * s.p.eta * s.p.dt / (w_total * s.p.dx3);
        tot[0] = w_total;
        tot[1] = epw;
    }
}

/* Enthalpy of a non-GAS cell as the thermal kernel reads it: after this step's arc injection
 * (reference kernels.py:497-510) and enthalpy ceiling (coupled_step.py:108). */
struct HView {
    const waam_step_fields *f;
    const StepConsts *s;
    ArcParams ap;
    float epw;
    bool arc_on, cap_on;

    __device__ __forceinline__ float at(size_t c, int i, int j, int k) const
    {
        float h = f->H_in[c];
        if (arc_on && i >= s->box_i0 && i <= s->box_i1 && j >= s->box_j0 && j <= s->box_j1 &&
            k >= s->box_k0 && k <= s->box_k1) {
            float wg, wdep;
            if (arc_cell_weights(ap, i, j, k, f->fl_in[c], wg, wdep)) h += epw * wg * wdep;
        }
        if (cap_on) h = ceiling_clamp(h, s->pc.H_liq, s->cap_l, s->cap_s);
        return h;
    }
    /* _masked_at on the enthalpy view: clamped index, GAS neighbours give the fallback */
    __device__ __forceinline__ float masked(int i, int j, int k, float fallback) const
    {
        const Dims &d = s->d;
        int ii = imax(0, imin(i, d.nx - 1));
        int jj = imax(0, imin(j, d.ny - 1));
        int kk = imax(0, imin(k, d.nz - 1));
        size_t c = d.idx(ii, jj, kk);
        if (f->flags_in[c] == GAS) return fallback;
        return at(c, ii, jj, kk);
    }
};

/* 3 — thermal chain (reference kernels.py:497-510, 1953-1984, 1611-1682, 2030-2082,
 *     1937-1950, 1519-1556, 3016-3032, 2013-2027 in coupled_step order) */
__global__ void __launch_bounds__(256) k_step_thermal(waam_step_fields f, StepConsts s)
{
    const Dims &d = s.d;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int i = blockIdx.z;
    if (k >= d.nz || j >= d.ny) return;
    const waam_step_params &p = s.p;
    size_t c = d.idx(i, j, k);
    uint8_t flag = f.flags_in[c];
    if (flag == GAS) {
        /* GAS: only the two phase updates touch it (ceiling, thermal and losses skip GAS) */
        float h = f.H_in[c], T, fl;
        phase_of(h, s.pc, T, fl);
        f.H_out[c] = h;
        f.T_out[c] = T;
        f.fl_out[c] = fl;
        return;
    }
    const float *tot = reinterpret_cast<const float *>((char *)f.scratch + kScratchTelemetry);
    HView hv;
    hv.f = &f;
    hv.s = &s;
    hv.ap = s.arc;
    hv.ap.arc_k = step_arc_k(p, tot);
    hv.epw = tot[1];
    hv.arc_on = p.is_welding && !p.arc_applied && hv.epw > 0.0f;
    hv.cap_on = p.is_welding && !p.arc_applied && p.enable_cap;

    /* advect_diffuse_temperature (Jacobi: every neighbour value is pre-kernel) */
    float T_c = f.T_in[c];
    float T_im = masked_at(f.T_in, f.flags_in, i - 1, j, k, T_c, d);
    float T_ip = masked_at(f.T_in, f.flags_in, i + 1, j, k, T_c, d);
    float T_jm = masked_at(f.T_in, f.flags_in, i, j - 1, k, T_c, d);
    float T_jp = masked_at(f.T_in, f.flags_in, i, j + 1, k, T_c, d);
    float T_km = masked_at(f.T_in, f.flags_in, i, j, k - 1, T_c, d);
    float T_kp = masked_at(f.T_in, f.flags_in, i, j, k + 1, T_c, d);
    float lap = T_ip + T_im + T_jp + T_jm + T_kp + T_km - 6.0f * T_c;
    float dH = p.cp_rho * p.alpha_lu * lap * 1.0f;
    float H_c = hv.at(c, i, j, k);
    if (flag != SOLID) {
        float u = f.ux[c], v = f.uy[c], w = f.uz[c];
        float H_im = hv.masked(i - 1, j, k, H_c);
        float H_ip = hv.masked(i + 1, j, k, H_c);
        float H_imm = hv.masked(i - 2, j, k, H_im);
        float H_ipp = hv.masked(i + 2, j, k, H_ip);
        float H_jm = hv.masked(i, j - 1, k, H_c);
        float H_jp = hv.masked(i, j + 1, k, H_c);
        float H_jmm = hv.masked(i, j - 2, k, H_jm);
        float H_jpp = hv.masked(i, j + 2, k, H_jp);
        float H_km = hv.masked(i, j, k - 1, H_c);
        float H_kp = hv.masked(i, j, k + 1, H_c);
        float H_kmm = hv.masked(i, j, k - 2, H_km);
        float H_kpp = hv.masked(i, j, k + 2, H_kp);
        float gx = upwind_grad(u, H_imm, H_im, H_c, H_ip, H_ipp);
        float gy = upwind_grad(v, H_jmm, H_jm, H_c, H_jp, H_jpp);
        float gz = upwind_grad(w, H_kmm, H_km, H_c, H_kp, H_kpp);
        dH -= (u * gx + v * gy + w * gz) * 1.0f;
    }
    float h = H_c + dH;
    if (p.enable_heat_loss) {
        int n_exp = exposed_faces(f.flags_in, i, j, k, d);
        if (n_exp != 0) h -= loss_flux(T_c, s.lc) * (float)n_exp * s.lc.dt / s.lc.dx;
        if (h < s.h_min) h = s.h_min;
    }
    if (p.enable_cap) h = ceiling_clamp(h, s.pc.H_liq, s.cap_l, s.cap_s);
    float T, fl;
    phase_of(h, s.pc, T, fl);
    f.H_out[c] = h;
    f.T_out[c] = T;
    f.fl_out[c] = fl;
    /* HAZ diagnostics (reference kernels.py:3027-3032, 2023-2027); they run before the VOF
     * flag update, so they see this step's incoming flags */
    if (T > f.T_max[c]) f.T_max[c] = T;
    f.dT_dt[c] = (T - f.T_prev[c]) * s.inv_dt;
    f.T_prev[c] = T;
}

/* update_time_above_T (reference kernels.py:3261-3283) runs after the VOF flag update and
 * skips cells that are GAS by the updated flags */
__device__ __forceinline__ void time_above(const waam_step_fields &f, const waam_step_params &p,
                                           size_t c)
{
    float T = f.T_out[c];
    if (T >= p.T800) f.t800[c] += p.dt;
    if (T >= p.T1100) f.t1100[c] += p.dt;
    if (T >= p.T_s) f.tsol[c] += p.dt;
}

/* 4 — advect_phi + reinitialize_phi (reference kernels.py:2320-2391) */
__global__ void __launch_bounds__(256) k_step_vof_advect(waam_step_fields f, Dims d)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int i = blockIdx.z;
    if (k >= d.nz || j >= d.ny) return;
    size_t c = d.idx(i, j, k);
    float p = advected_phi(f.phi, f.ux, f.uy, f.uz, f.flags_in, i, j, k, d);
    f.phi_tmp[c] = reinit_phi(p, f.flags_in[c]);
}

/* 5 — apply_contact_angle_phi_bc (Jacobi), update_flags_from_phi, sync_phi_liquid_fraction,
 *     update_time_above_T (reference kernels.py:933-991, 2394-2419, 3484-3504, 3261-3283) */
__global__ void __launch_bounds__(256) k_step_vof_flags(waam_step_fields f, StepConsts s)
{
    const Dims &d = s.d;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int i = blockIdx.z;
    if (k >= d.nz || j >= d.ny) return;
    size_t c = d.idx(i, j, k);
    if (!s.p.enable_vof) {
        uint8_t fi = f.flags_in[c];
        f.flags_out[c] = fi;
        if (fi != GAS) time_above(f, s.p, c);
        return;
    }
    float p = s.p.enable_wetting
                  ? contact_bc_phi(f.phi_tmp, f.flags_in, i, j, k, s.p.cos_t, s.p.sin_t, d)
                  : f.phi_tmp[c];
    uint8_t nf = classify_cell(k, s.p.nz_solid, p, f.fl_out[c]);
    f.phi[c] = p;
    f.flags_out[c] = nf;
    if (nf == GAS || p < 0.05f) f.fl_out[c] = 0.0f;
    else if (nf == SOLID) f.fl_out[c] = 0.0f;
    if (nf != GAS) time_above(f, s.p, c);
}

/* pull of the reference stream (kernels.py:2991-3010) with the flags of the step it belongs to */
__device__ __forceinline__ void pull_cell(const float *__restrict__ src, float *__restrict__ dst,
                                          const uint8_t *__restrict__ sfl, int i, int j, int k,
                                          uint8_t fc, const Dims &d)
{
    size_t N = d.n();
    size_t c = d.idx(i, j, k);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        int ni = i - ex(q), nj = j - ey(q), nk = k - ez(q);
        ni = ni < 0 ? ni + d.nx : (ni >= d.nx ? ni - d.nx : ni);
        nj = imax(0, imin(nj, d.ny - 1));
        nk = imax(0, imin(nk, d.nz - 1));
        size_t n = d.idx(ni, nj, nk);
        float v;
        if (fc == GAS) v = src[(size_t)q * N + n];
        else if (sfl[n] == SOLID) v = src[(size_t)opp(q) * N + c];
        else v = src[(size_t)q * N + n];
        dst[(size_t)q * N + c] = v;
    }
}

/* 6 — body forces, collision, Mach clamp; pending pull stream of the previous step */
__global__ void __launch_bounds__(256) k_step_collide(waam_step_fields f, StepConsts s)
{
    const Dims &d = s.d;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int i = blockIdx.z;
    if (k >= d.nz || j >= d.ny) return;
    const waam_step_params &p = s.p;
    size_t c = d.idx(i, j, k);
    size_t N = d.n();
    uint8_t fc = f.flags_out[c];
    f.sflags_next[c] = fc;
    if (fc == SOLID || fc == GAS) {
        if (p.store_forces && !p.forces_from_fields) {
            f.Fx[c] = 0.0f;
            f.Fy[c] = 0.0f;
            f.Fz[c] = 0.0f;
        }
        if (p.pending_stream) {
            uint8_t fp = f.sflags_prev[c];
            if (fp != SOLID) pull_cell(f.f_src, f.f_dst, f.sflags_prev, i, j, k, fp, d);
        }
        return;
    }
    /* forces, reference order coupled_step.py:298-362 */
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if (p.forces_from_fields) {
        fx = f.Fx[c];
        fy = f.Fy[c];
        fz = f.Fz[c];
    }
    float fl = f.fl_out[c];
    float tx, ty, tz;
    if (p.enable_csf &&
        csf_force(f.phi, f.flags_out, i, j, k, d, p.gamma_lu, p.enable_wetting, p.cos_t, p.sin_t,
                  tx, ty, tz)) {
        fx += tx;
        fy += ty;
        fz += tz;
    }
    if (marangoni_force(f.T_out, f.phi, fl, i, j, k, d, p.dgamma, tx, ty, tz)) {
        fx += tx;
        fy += ty;
        fz += tz;
    }
    if (p.is_welding) {
        ArcPressure ap = s.ap;
        const float *tot = reinterpret_cast<const float *>((char *)f.scratch + kScratchTelemetry);
        ap.arc_k = step_arc_k(p, tot);
        if (arc_pressure_force(f.phi, i, j, k, d, ap, tz)) fz += tz;
    }
    if (p.enable_hydro && fl > 0.0f) fz += -1.0f * p.g_lu * fl;
    if (fl > 0.0f) fz += 1.0f * p.g_lu * p.beta * (f.T_out[c] - p.T_l) * fl;
    if (p.F_cap > 0.0f) {
        float mag = sqrtf(fx * fx + fy * fy + fz * fz);
        if (mag > p.F_cap) {
            float sc = p.F_cap / (mag + 1e-12f);
            fx = fx * sc;
            fy = fy * sc;
            fz = fz * sc;
        }
    }
    if (p.store_forces) {
        f.Fx[c] = fx;
        f.Fy[c] = fy;
        f.Fz[c] = fz;
    }
    float fin[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) fin[q] = f.f_src[(size_t)q * N + c];
    CellOut o;
    if (p.collision == 0) mrt_cell(fin, fx, fy, fz, fl, p.omega_s, p.omega_b, p.C_darcy, o);
    else srt_cell(fin, fx, fy, fz, fl, p.omega_s, p.C_darcy, o);
#pragma unroll
    for (int q = 0; q < Q; ++q) f.f_dst[(size_t)q * N + c] = o.f[q];
    float u = o.u, v = o.v, w = o.w;
    if (p.u_cap > 0.0f) {
        float mag = sqrtf(u * u + v * v + w * w);
        if (mag > p.u_cap) {
            float sc = p.u_cap / (mag + 1e-12f);
            u = u * sc;
            v = v * sc;
            w = w * sc;
        }
    }
    f.rho[c] = o.rho;
    f.ux[c] = u;
    f.uy[c] = v;
    f.uz[c] = w;
}

/* 8 — remelt_hot_solid_scalar, solidify_trailing_pool_scalar, solidify_cooled_metal
 *     (reference kernels.py:2489-2520, 2573-2617, 2422-2450), in that order per cell */
__global__ void __launch_bounds__(256) k_step_freeze(waam_step_fields f, StepConsts s)
{
    const Dims &d = s.d;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int i = blockIdx.z;
    if (k >= d.nz || j >= d.ny) return;
    const waam_step_params &p = s.p;
    size_t c = d.idx(i, j, k);
    uint8_t flag = f.flags_out[c];
    if (flag == GAS) return;
    float T = f.T_out[c];
    if (flag == SOLID && T > p.T_s + 5.0f) {
        float H_sol = p.cp_rho * p.T_s, fl;
        if (T >= p.T_l) {
            fl = 1.0f;
            f.H_out[c] = p.cp_rho * p.T_l + p.L_rho + p.cp_rho * (T - p.T_l);
        } else {
            fl = (T - p.T_s) / (p.T_l - p.T_s + 1e-6f);
            f.H_out[c] = H_sol + fl * p.L_rho;
        }
        f.fl_out[c] = fl;
        f.phi[c] = 1.0f;
        flag = FLUID;
        f.flags_out[c] = FLUID;
    }
    if (p.enable_trailing) {
        float phi = f.phi[c], fl = f.fl_out[c];
        if (!(phi < 0.85f || fl < 0.05f)) {
            const float *tot = reinterpret_cast<const float *>((char *)f.scratch + kScratchTelemetry);
            float ak = step_arc_k(p, tot);
            float di = (float)i - p.arc_i, dj = (float)j - p.arc_j, dk = (float)k - ak;
            float proj = di * p.dir_x + dj * p.dir_y + dk * p.dir_z;
            if (!(proj > -p.lookback) && !(T >= p.T_freeze)) {
                f.H_out[c] = p.cp_rho * (p.T_s - 1.0f);
                T = p.T_s - 1.0f;
                f.T_out[c] = T;
                f.fl_out[c] = 0.0f;
                f.phi[c] = 1.0f;
                flag = SOLID;
                f.flags_out[c] = SOLID;
                f.ux[c] = 0.0f;
                f.uy[c] = 0.0f;
                f.uz[c] = 0.0f;
            }
        }
    }
    if (f.phi[c] < 0.85f) return;
    if (f.fl_out[c] < 0.02f && T < p.T_s) {
        f.flags_out[c] = SOLID;
        f.phi[c] = 1.0f;
        f.fl_out[c] = 0.0f;
        if (p.zero_velocity == 1) {
            f.ux[c] = 0.0f;
            f.uy[c] = 0.0f;
            f.uz[c] = 0.0f;
        }
    }
}

/* standalone pending stream */
__global__ void k_stream_pending(const float *__restrict__ src, float *__restrict__ dst,
                                 const uint8_t *__restrict__ sfl, Dims d)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int i = blockIdx.z;
    if (k >= d.nz || j >= d.ny) return;
    uint8_t fp = sfl[d.idx(i, j, k)];
    if (fp != SOLID) pull_cell(src, dst, sfl, i, j, k, fp, d);
}

__global__ void k_advect_tracers_step(float *__restrict__ pos, int32_t *__restrict__ active,
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
    uint8_t fg = flags[c];
    if (fg == GAS) {
        active[p] = 0;
        return;
    }
    if (f_l[c] < 0.05f || fg == SOLID) {
        active[p] = 2;
        return;
    }
    pos[3 * p] += ux[c] * dx;
    pos[3 * p + 1] += uy[c] * dx;
    pos[3 * p + 2] += uz[c] * dx;
}

static void arc_box(StepConsts &s)
{
    const waam_step_params &p = s.p;
    /* exp(x) is exactly zero for x < -103.98; outside this box x < -110 for every cell */
    double ri, rj, rk;
    if (p.arc_kind == 0) {
        double inv = (double)s.arc.inv2s2;
        ri = rj = std::sqrt(110.0 / inv);
        rk = 2.0 * ri;
    } else {
        double a = std::max((double)p.g_af, (double)p.g_ar);
        ri = std::sqrt(110.0 / 3.0 * (a * a + 1e-6)) / std::max(std::fabs((double)p.g_sgn), 1e-30);
        rj = std::sqrt(110.0 / 3.0 * ((double)p.g_b * p.g_b + 1e-6));
        rk = std::sqrt(110.0 / 3.0 * ((double)p.g_c * p.g_c + 1e-6));
    }
    auto lo = [](double a, double r) { return (int)std::floor(a - r) - 2; };
    auto hi = [](double a, double r) { return (int)std::ceil(a + r) + 2; };
    s.box_i0 = lo(p.arc_i, ri);
    s.box_i1 = hi(p.arc_i, ri);
    s.box_j0 = lo(p.arc_j, rj);
    s.box_j1 = hi(p.arc_j, rj);
    /* arc_k is only known on the device; bound it by the grid (k spans the whole column) */
    (void)rk;
    s.box_k0 = -1;
    s.box_k1 = s.d.nz + 1;
    if (!(ri < 1e6) || !(rj < 1e6)) {
        s.box_i0 = s.box_j0 = -(1 << 30);
        s.box_i1 = s.box_j1 = (1 << 30);
    }
}

/* HIP-event timing of one kernel kind (waam_timing_begin / waam_timing_end) */
struct KernelTimer {
    bool armed = false;
    int which = -1, used = 0;
    std::vector<hipEvent_t> ev;  // start/stop pairs
};
static KernelTimer g_timer;

template <class Launch>
static void timed_launch(int which, hipStream_t st, Launch &&launch)
{
    KernelTimer &t = g_timer;
    bool rec = t.armed && t.which == which && 2 * t.used < (int)t.ev.size();
    if (rec) hipEventRecord(t.ev[2 * t.used], st);
    launch();
    if (rec) {
        hipEventRecord(t.ev[2 * t.used + 1], st);
        ++t.used;
    }
}

}  // namespace waam

using namespace waam;

extern "C" {

int waam_step_fused(const waam_step_params *pp, const waam_step_fields *ff, void *stream)
{
    if (!pp || !ff) return set_error(WAAM_ERR_INVALID, "waam_step_fused: NULL argument");
    const waam_step_params &p = *pp;
    const waam_step_fields &f = *ff;
    WAAM_RETURN_IF(check_dims("waam_step_fused", f.nx, f.ny, f.nz));
    WAAM_RETURN_IF(check_ptrs("waam_step_fused",
                              {f.f_src, f.f_dst, f.rho, f.ux, f.uy, f.uz, f.H_in, f.H_out, f.T_in,
                               f.T_out, f.fl_in, f.fl_out, f.phi, f.phi_tmp, f.flags_in,
                               f.flags_out, f.sflags_prev, f.sflags_next, f.Fx, f.Fy, f.Fz,
                               f.T_max, f.T_prev, f.dT_dt, f.t800, f.t1100, f.tsol, f.scratch}));
    if (f.f_src == f.f_dst || f.H_in == f.H_out || f.T_in == f.T_out || f.fl_in == f.fl_out ||
        f.flags_in == f.flags_out || f.sflags_prev == f.sflags_next || f.phi == f.phi_tmp)
        return set_error(WAAM_ERR_INVALID, "waam_step_fused: an in/out buffer pair aliases");
    if (p.col_i < 0 || p.col_i >= f.nx || p.col_j < 0 || p.col_j >= f.ny)
        return set_error(WAAM_ERR_INVALID, "waam_step_fused: arc column outside the grid");
    if (f.max_tracers > 0 && (!f.tracer_pos || !f.tracer_active))
        return set_error(WAAM_ERR_INVALID, "waam_step_fused: tracer buffers missing");

    StepConsts s{};
    s.p = p;
    s.d = Dims{f.nx, f.ny, f.nz};
    s.arc.kind = p.arc_kind;
    s.arc.arc_i = p.arc_i;
    s.arc.arc_j = p.arc_j;
    s.arc.inv2s2 = 1.0f / (2.0f * p.sigma * p.sigma);
    s.arc.sgn = p.g_sgn;
    s.arc.ff = p.g_ff;
    s.arc.fr = p.g_fr;
    s.arc.af = p.g_af;
    s.arc.ar = p.g_ar;
    s.arc.b = p.g_b;
    s.arc.cax = p.g_c;
    s.arc.pen = p.pen;
    s.arc.enable = p.surface_weight;
    s.pc = make_phase_consts(p.cp_rho, p.L_rho, p.T_s, p.T_l);
    s.lc = LossConsts{p.T_amb, p.h_conv, p.eps_rad, p.sigma_sb, p.dt, p.dx, p.enable_conv, p.enable_rad};
    s.cap_l = s.pc.H_liq + p.cp_rho * (p.T_cap - p.T_l);
    s.cap_s = p.cp_rho * p.T_cap;
    s.h_min = p.cp_rho * p.T_amb;
    s.inv_dt = 1.0f / (p.dt + 1e-12f);
    s.ap.arc_i = p.arc_i;
    s.ap.arc_j = p.arc_j;
    s.ap.inv2s2 = 1.0f / (2.0f * p.p_sigma * p.p_sigma + 1e-6f);
    s.ap.F_peak_lu = p.p_pa / (p.rho_phys * p.dx) * (p.dt * p.dt) / p.dx;
    arc_box(s);

    hipStream_t st = as_stream(stream);
    const Dims &d = s.d;
    bool arc_sums = p.is_welding && !p.arc_applied;
    if (!p.use_arc_k_override || arc_sums)
        timed_launch(WAAM_K_ARC_PARTIALS, st, [&] {
            hipLaunchKernelGGL(k_step_arc_partials, dim3(arc_sums ? kReduceBlocks : 1),
                               dim3(kBlock), 0, st, f, s, arc_sums ? 1 : 0);
        });
    if (arc_sums)
        timed_launch(WAAM_K_ARC_TOTAL, st, [&] {
            hipLaunchKernelGGL(k_step_arc_total, dim3(1), dim3(kBlock), 0, st, f, s);
        });
    timed_launch(WAAM_K_THERMAL, st,
                 [&] { hipLaunchKernelGGL(k_step_thermal, grid3(d), block3(), 0, st, f, s); });
    if (p.enable_vof)
        timed_launch(WAAM_K_VOF_ADVECT, st, [&] {
            hipLaunchKernelGGL(k_step_vof_advect, grid3(d), block3(), 0, st, f, d);
        });
    timed_launch(WAAM_K_VOF_FLAGS, st,
                 [&] { hipLaunchKernelGGL(k_step_vof_flags, grid3(d), block3(), 0, st, f, s); });
    timed_launch(WAAM_K_COLLIDE, st,
                 [&] { hipLaunchKernelGGL(k_step_collide, grid3(d), block3(), 0, st, f, s); });
    if (f.max_tracers > 0)
        timed_launch(WAAM_K_TRACERS, st, [&] {
            hipLaunchKernelGGL(k_advect_tracers_step, dim3(grid_for((size_t)f.max_tracers)),
                               dim3(kBlock), 0, st, f.tracer_pos, f.tracer_active, f.ux, f.uy,
                               f.uz, f.fl_out, f.flags_out, p.dx, f.max_tracers, d);
        });
    if (p.enable_freeze)
        timed_launch(WAAM_K_FREEZE, st, [&] {
            hipLaunchKernelGGL(k_step_freeze, grid3(d), block3(), 0, st, f, s);
        });
    return check_launch("waam_step_fused");
}

int waam_stream_pending(const float *f_src, float *f_dst, const uint8_t *sflags, int nx, int ny,
                        int nz, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_stream_pending", nx, ny, nz));
    WAAM_RETURN_IF(check_ptrs("waam_stream_pending", {f_src, f_dst, sflags}));
    if (f_src == f_dst) return set_error(WAAM_ERR_INVALID, "waam_stream_pending: f_src aliases f_dst");
    Dims d{nx, ny, nz};
    hipLaunchKernelGGL(k_stream_pending, grid3(d), block3(), 0, as_stream(stream), f_src, f_dst,
                       sflags, d);
    return check_launch("waam_stream_pending");
}

int waam_step_arc_readback(const void *scratch, float *host_out3)
{
    if (!scratch || !host_out3) return set_error(WAAM_ERR_INVALID, "waam_step_arc_readback: NULL");
    hipError_t e = hipMemcpy(host_out3, (const char *)scratch + kScratchTelemetry, 3 * sizeof(float),
                             hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_error(WAAM_ERR_HIP, "waam_step_arc_readback: %s", hipGetErrorString(e));
    return WAAM_OK;
}

int waam_timing_begin(int which, int max_launches)
{
    KernelTimer &t = g_timer;
    if (t.armed) return set_error(WAAM_ERR_INVALID, "waam_timing_begin: timing already armed");
    if (which < 0 || which >= WAAM_K_COUNT)
        return set_error(WAAM_ERR_INVALID, "waam_timing_begin: kernel id %d out of range", which);
    if (max_launches <= 0 || max_launches > (1 << 20))
        return set_error(WAAM_ERR_INVALID, "waam_timing_begin: max_launches %d out of range",
                         max_launches);
    t.ev.assign((size_t)2 * max_launches, nullptr);
    for (auto &e : t.ev) {
        hipError_t err = hipEventCreate(&e);
        if (err != hipSuccess) {
            for (auto &x : t.ev)
                if (x) hipEventDestroy(x);
            t.ev.clear();
            return set_error(WAAM_ERR_HIP, "waam_timing_begin: %s", hipGetErrorString(err));
        }
    }
    t.which = which;
    t.used = 0;
    t.armed = true;
    return WAAM_OK;
}

int waam_timing_end(double *total_ms, int64_t *launches)
{
    KernelTimer &t = g_timer;
    if (!t.armed) return set_error(WAAM_ERR_INVALID, "waam_timing_end: timing not armed");
    double sum = 0.0;
    int rc = WAAM_OK;
    for (int i = 0; i < t.used; ++i) {
        float ms = 0.0f;
        hipError_t e = hipEventSynchronize(t.ev[2 * i + 1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t.ev[2 * i], t.ev[2 * i + 1]);
        if (e != hipSuccess) {
            rc = set_error(WAAM_ERR_HIP, "waam_timing_end: %s", hipGetErrorString(e));
            break;
        }
        sum += ms;
    }
    for (auto &e : t.ev) hipEventDestroy(e);
    t.ev.clear();
    t.armed = false;
    if (total_ms) *total_ms = sum;
    if (launches) *launches = t.used;
    return rc;
}

}  // extern "C"
