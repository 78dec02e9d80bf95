// adjoint.hip — reverse-mode derivative of the thermal sub-step (the differentiation form).
//
// The forward map is the coupled step's thermal chain with frozen flags and zero velocity
// (reference solvers/coupled_step.py:97-160 with the flow and the free surface held): for every
// non-GAS cell
//   A = H_n + dH_arc            arc injection (Gaussian2D, surface weighting), if welding
//   B = ceil(A)                 enthalpy ceiling, if enabled
//   C = B + k lap(T_n)          conduction from the step's starting temperature
//   E = C - n_exp dt/dx q(T_n)  boundary losses (convection + radiation), if enabled
//   F = max(E, H_floor)         enthalpy floor, with the losses
//   H_{n+1} = ceil(F)           (phase map T, f_l = phase(H) after each)
// with T_n = phase(H_n). Its parameters are the arc efficiency eta, a factor s_h on both loss
// coefficients and a factor s_sigma on the arc radius. Given lambda_{n+1} = dL/dH_{n+1}, one
// adjoint step returns lambda_n and adds this step's terms of dL/d(eta, s_h, s_sigma) to three
// binary64 accumulators (block partial sums combined in a fixed order: deterministic).
// Derivatives at the piecewise points are one-sided, taken on the branch the forward took
// (a clamped value has derivative 0; the phase map's slope is that of the side the enthalpy
// is on). The surface weight's dependence on f_l (a step at f_l = 0.55) has derivative 0.
#include "arc_cell.h"
#include "thermal_cell.h"

namespace waam {

/* blocks of the adjoint's reductions: 6 doubles each fit the workspace's telemetry part */
constexpr int kAdjBlocks = 1024;
static_assert(6 * kAdjBlocks <= 8 * kReduceBlocks, "adjoint partials exceed the workspace");

struct AdjC {
    PhaseConsts pc;
    float cap_l, cap_s, h_floor, kcond, loss_geo;
    LossConsts lc0;  // unscaled loss coefficients
    float s_h;
    ArcParams ap;
    float sigma, sigma0, K;  // K = Q eta dt / dx3
    int enable_cap, enable_loss, enable_arc;
    float eta;
};

static AdjC make_adj(const waam_adj_consts &a)
{
    AdjC c{};
    c.pc = make_phase_consts(a.cp_rho, a.L_rho, a.T_s, a.T_l);
    c.cap_l = c.pc.H_liq + a.cp_rho * (a.T_cap - a.T_l);
    c.cap_s = a.cp_rho * a.T_cap;
    c.h_floor = a.cp_rho * a.T_floor;
    c.kcond = a.cp_rho * a.alpha_lu * 1.0f;
    c.loss_geo = a.dt / a.dx;
    c.lc0 = LossConsts{a.T_amb, a.h_conv0, a.eps_rad0, a.sigma_sb, a.dt, a.dx, a.enable_conv,
                       a.enable_rad};
    c.s_h = a.loss_scale;
    c.ap = ArcParams{};
    c.ap.kind = 0;
    c.ap.arc_i = a.arc_i;
    c.ap.arc_j = a.arc_j;
    c.ap.arc_k = a.arc_k;
    c.ap.inv2s2 = 1.0f / (2.0f * a.sigma * a.sigma);
    c.ap.pen = a.pen;
    c.ap.enable = a.surface_weight;
    c.sigma = a.sigma;
    c.sigma0 = a.sigma0;
    c.K = a.Q_w * a.eta * a.dt / a.dx3;
    c.enable_cap = a.enable_cap;
    c.enable_loss = a.enable_loss;
    c.enable_arc = a.enable_arc;
    c.eta = a.eta;
    return c;
}

__device__ __forceinline__ void phase_slopes(float h, const PhaseConsts &p, float &dT, float &dfl)
{
    if (h <= p.H_sol) {
        dT = 1.0f / p.cp_rho;
        dfl = 0.0f;
    } else if (h >= p.H_liq) {
        dT = 1.0f / p.cp_rho;
        dfl = 0.0f;
    } else {
        dfl = 1.0f / p.L_rho;
        dT = (p.T_l - p.T_s) / p.L_rho;
    }
}

__device__ __forceinline__ bool ceiling_clamps(float h, const AdjC &a)
{
    return h > a.pc.H_liq ? h > a.cap_l : h > a.cap_s;
}

/* dq/dT of the loss flux (q is evaluated at min(T, 2500)) */
__device__ __forceinline__ float loss_slope(float T, const LossConsts &lc)
{
    if (!(T < 2500.0f)) return 0.0f;
    float s = 0.0f;
    if (lc.enable_conv == 1) s += lc.h_conv;
    if (lc.enable_rad == 1) s += 4.0f * lc.eps_rad * lc.sigma_sb * (T * T * T);
    return s;
}

/* W = sum wg wdep and S = sum wg wdep r^2 over the non-GAS cells the arc weights */
__global__ void __launch_bounds__(kBlock) k_adj_arc_partials(const uint8_t *__restrict__ flags,
                                                             const float *__restrict__ fl,
                                                             AdjC a, Dims d,
                                                             double *__restrict__ parts)
{
    __shared__ double sh_s[kBlock], sh_c[kBlock];
    double w = 0.0, wc = 0.0, s = 0.0, sc = 0.0;
    size_t N = d.n();
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N;
         c += (size_t)gridDim.x * blockDim.x) {
        if (flags[c] == GAS) continue;
        int i, j, k;
        d.ijk(c, i, j, k);
        float wg, wdep;
        if (!arc_cell_weights(a.ap, i, j, k, fl[c], wg, wdep)) continue;
        float di = (float)i - a.ap.arc_i, dj = (float)j - a.ap.arc_j, dk = (float)k - a.ap.arc_k;
        float r2 = di * di + dj * dj + dk * dk * 0.25f;
        double x = (double)(wg * wdep), ss, e;
        two_sum(w, x, ss, e);
        w = ss;
        wc += e;
        two_sum(s, x * (double)r2, ss, e);
        s = ss;
        sc += e;
    }
    block_reduce_pair(w, wc, sh_s, sh_c);
    if (threadIdx.x == 0) {
        parts[4 * blockIdx.x] = w;
        parts[4 * blockIdx.x + 1] = wc;
    }
    __syncthreads();
    block_reduce_pair(s, sc, sh_s, sh_c);
    if (threadIdx.x == 0) {
        parts[4 * blockIdx.x + 2] = s;
        parts[4 * blockIdx.x + 3] = sc;
    }
}

/* fixed-order total of npart groups of `width` compensated pairs: out[g] (+)= sum */
__global__ void __launch_bounds__(kBlock) k_adj_total(const double *__restrict__ parts, int npart,
                                                      int width, double *__restrict__ out,
                                                      int accumulate)
{
    __shared__ double sh_s[kBlock], sh_c[kBlock];
    for (int g = 0; g < width; ++g) {
        double s = 0.0, cmp = 0.0;
        for (int b = threadIdx.x; b < npart; b += blockDim.x) {
            double ss, e;
            two_sum(s, parts[(size_t)b * 2 * width + 2 * g], ss, e);
            s = ss;
            cmp += parts[(size_t)b * 2 * width + 2 * g + 1] + e;
        }
        block_reduce_pair(s, cmp, sh_s, sh_c);
        if (threadIdx.x == 0) out[g] = (accumulate ? out[g] : 0.0) + (s + cmp);
        __syncthreads();
    }
}

/* lambda = dL/dT dT/dH + dL/df_l df_l/dH at the final enthalpy (0 on GAS) */
__global__ void k_adj_terminal(const float *__restrict__ H, const float *__restrict__ gT,
                               const float *__restrict__ gfl, const uint8_t *__restrict__ flags,
                               PhaseConsts pc, float *__restrict__ lam, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    if (flags[c] == GAS) {
        lam[c] = 0.0f;
        return;
    }
    float dT, dfl;
    phase_slopes(H[c], pc, dT, dfl);
    lam[c] = gT[c] * dT + gfl[c] * dfl;
}

/* adjoint step, part 1: through the second ceiling and the floor; lamC = dL/dC (0 on GAS) */
__global__ void k_adj_step1(const float *__restrict__ lam_next, const float *__restrict__ E,
                            const float *__restrict__ F, const uint8_t *__restrict__ flags, AdjC a,
                            float *__restrict__ lamC, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    if (flags[c] == GAS) {
        lamC[c] = 0.0f;
        return;
    }
    float l = lam_next[c];
    if (a.enable_cap && ceiling_clamps(F[c], a)) l = 0.0f;
    if (a.enable_loss && E[c] < a.h_floor) l = 0.0f;
    lamC[c] = l;
}

/* adjoint step, part 2: losses, conduction (the symmetric 6-neighbour operator over in-range
 * non-GAS neighbours), first ceiling, arc; lam_out = dL/dH_n; per-block partial sums of
 * (d/d eta, d/d s_h, d/d s_sigma) */
__global__ void __launch_bounds__(kBlock) k_adj_step2(
    const float *__restrict__ lamC, const float *__restrict__ H_n, const float *__restrict__ A,
    const float *__restrict__ fl_n, const uint8_t *__restrict__ flags,
    const double *__restrict__ arc_sums, AdjC a, Dims d, float *__restrict__ lam_out,
    double *__restrict__ parts)
{
    __shared__ double sh_s[kBlock], sh_c[kBlock];
    double g[3] = {0.0, 0.0, 0.0}, gc[3] = {0.0, 0.0, 0.0};
    size_t N = d.n();
    const double W = a.enable_arc ? arc_sums[0] : 0.0;
    const double S = a.enable_arc ? arc_sums[1] : 0.0;
    const float epw = (a.enable_arc && (float)W > 1e-9f) ? a.K / (float)W : 0.0f;
    const float rbar = W > 0.0 ? (float)(S / W) : 0.0f;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N;
         c += (size_t)gridDim.x * blockDim.x) {
        if (flags[c] == GAS) {
            lam_out[c] = 0.0f;
            continue;
        }
        int i, j, k;
        d.ijk(c, i, j, k);
        float lc = lamC[c];
        float h = H_n[c];
        float Tn, fln;
        phase_of(h, a.pc, Tn, fln);
        /* conduction transpose: the same stencil on lamC */
        float acc = 0.0f;
        const int di[6] = {-1, 1, 0, 0, 0, 0}, dj[6] = {0, 0, -1, 1, 0, 0},
                  dk[6] = {0, 0, 0, 0, -1, 1};
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            int ii = i + di[t], jj = j + dj[t], kk = k + dk[t];
            if (ii < 0 || ii >= d.nx || jj < 0 || jj >= d.ny || kk < 0 || kk >= d.nz) continue;
            size_t n = d.idx(ii, jj, kk);
            if (flags[n] == GAS) continue;
            acc += lamC[n] - lc;
        }
        float lT = a.kcond * acc;
        if (a.enable_loss) {
            int n_exp = exposed_faces(flags, i, j, k, d);
            if (n_exp > 0) {
                LossConsts ls = a.lc0;
                ls.h_conv = a.lc0.h_conv * a.s_h;
                ls.eps_rad = a.lc0.eps_rad * a.s_h;
                float geo = (float)n_exp * a.loss_geo;
                lT -= lc * geo * loss_slope(Tn, ls);
                double x = (double)(-lc * geo * loss_flux(Tn, a.lc0)), ss, e;
                two_sum(g[1], x, ss, e);
                g[1] = ss;
                gc[1] += e;
            }
        }
        float lA = (a.enable_cap && ceiling_clamps(A[c], a)) ? 0.0f : lc;
        float dT, dfl;
        phase_slopes(h, a.pc, dT, dfl);
        lam_out[c] = lA + lT * dT;
        if (a.enable_arc && epw > 0.0f && lA != 0.0f) {
            float wg, wdep;
            if (arc_cell_weights(a.ap, i, j, k, fl_n[c], wg, wdep)) {
                float fi = (float)i - a.ap.arc_i, fj = (float)j - a.ap.arc_j,
                      fk = (float)k - a.ap.arc_k;
                float r2 = fi * fi + fj * fj + fk * fk * 0.25f;
                double dh = (double)epw * (double)wg * (double)wdep;
                double ss, e;
                two_sum(g[0], (double)lA * dh / (double)a.eta, ss, e);
                g[0] = ss;
                gc[0] += e;
                double s3 = (double)a.sigma * a.sigma * a.sigma;
                two_sum(g[2], (double)lA * dh * (double)a.sigma0 * ((double)r2 - rbar) / s3, ss,
                        e);
                g[2] = ss;
                gc[2] += e;
            }
        }
    }
    for (int t = 0; t < 3; ++t) {
        double s = g[t], cc = gc[t];
        block_reduce_pair(s, cc, sh_s, sh_c);
        if (threadIdx.x == 0) {
            parts[6 * blockIdx.x + 2 * t] = s;
            parts[6 * blockIdx.x + 2 * t + 1] = cc;
        }
        __syncthreads();
    }
}

}  // namespace waam

using namespace waam;

/* Adjoint of compute_marangoni_force (reference kernels.py:2158-2217) with phi, f_l and the
 * flags held: F_c = dgamma * gm_c * (I - n_c n_c^T) grad T_c on the cells it applies to, linear
 * in T. Pass 1, per cell c: G_c = dgamma * gm_c * (I - n n^T) lambda_c (the projection is
 * symmetric) and s_c = lambda_c . F_c / dgamma, both in binary64 from the forward's fp32 n and
 * gm; zero where the forward adds nothing. */
__global__ void k_adj_marangoni_project(const float *__restrict__ T, const float *__restrict__ phi,
                                        const float *__restrict__ f_l,
                                        const uint8_t *__restrict__ flags,
                                        const float *__restrict__ lx, const float *__restrict__ ly,
                                        const float *__restrict__ lz, float dgamma,
                                        double *__restrict__ G, double *__restrict__ s_cell,
                                        Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    const size_t N = d.n();
    const size_t c = d.idx(i, j, k);
    double gx = 0.0, gy = 0.0, gz = 0.0, sc = 0.0;
    const uint8_t fg = flags[c];
    if (fg != SOLID && fg != GAS && !(f_l[c] < 0.05f)) {
        const int ip = imin(i + 1, d.nx - 1), im = imax(i - 1, 0);
        const int jp = imin(j + 1, d.ny - 1), jm = imax(j - 1, 0);
        const int kp = imin(k + 1, d.nz - 1), km = imax(k - 1, 0);
        float dpx = 0.5f * (phi[d.idx(ip, j, k)] - phi[d.idx(im, j, k)]);
        float dpy = 0.5f * (phi[d.idx(i, jp, k)] - phi[d.idx(i, jm, k)]);
        float dpz = 0.5f * (phi[d.idx(i, j, kp)] - phi[d.idx(i, j, km)]);
        float gm = sqrtf(dpx * dpx + dpy * dpy + dpz * dpz);
        if (!(gm < 1e-6f)) {
            const double nx = dpx / gm, ny = dpy / gm, nz = dpz / gm;
            const double ax = lx[c], ay = ly[c], az = lz[c];
            const double nl = nx * ax + ny * ay + nz * az;
            const double w = (double)dgamma * (double)gm;
            gx = w * (ax - nl * nx);
            gy = w * (ay - nl * ny);
            gz = w * (az - nl * nz);
            const double tx = 0.5 * ((double)T[d.idx(ip, j, k)] - (double)T[d.idx(im, j, k)]);
            const double ty = 0.5 * ((double)T[d.idx(i, jp, k)] - (double)T[d.idx(i, jm, k)]);
            const double tz = 0.5 * ((double)T[d.idx(i, j, kp)] - (double)T[d.idx(i, j, km)]);
            const double nt = nx * tx + ny * ty + nz * tz;
            sc = (double)gm * (ax * (tx - nt * nx) + ay * (ty - nt * ny) + az * (tz - nt * nz));
        }
    }
    G[c] = gx;
    G[N + c] = gy;
    G[2 * N + c] = gz;
    s_cell[c] = sc;
}

/* Pass 2, gather: dL/dT at m += 0.5 * (G of the cells whose clamped +/- neighbour is m, with
 * the sign of that side). Along x: +G_x of m-1 (i >= 1) and of m itself at i = nx-1, -G_x of
 * m+1 (i <= nx-2) and of m itself at i = 0; likewise along y and z. */
__global__ void k_adj_marangoni_gather(const double *__restrict__ G, double *__restrict__ T_adj,
                                       Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    const size_t N = d.n();
    const size_t c = d.idx(i, j, k);
    const int n[3] = {d.nx, d.ny, d.nz};
    const int at[3] = {i, j, k};
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double *Ga = G + (size_t)a * N;
        const int v = at[a];
        int lo[3] = {i, j, k}, hi[3] = {i, j, k};
        lo[a] = v - 1;
        hi[a] = v + 1;
        if (v >= 1) acc += Ga[d.idx(lo[0], lo[1], lo[2])];
        if (v == n[a] - 1) acc += Ga[c];
        if (v <= n[a] - 2) acc -= Ga[d.idx(hi[0], hi[1], hi[2])];
        if (v == 0) acc -= Ga[c];
    }
    T_adj[c] += 0.5 * acc;
}

/* Adjoint of collide_srt (reference kernels.py:2815-2885: BGK with Guo forcing and the
 * Carman-Kozeny drag) with f_l and the flags held, on the cells it updates (neither SOLID nor
 * GAS; elsewhere it writes nothing, so those cells' derivatives are 0). For the adjoint lam of
 * the post-collision populations: lam_f = dL/df (19 q-major arrays) and lam_F = dL/dF (3 arrays),
 * in binary64 from the fp32 inputs. With r = sum f, u = (j + F/2) / ((r + eps) ck):
 *   out_q = (1 - omega) f_q + omega feq_q(r, u) + S_q(u, F),
 *   feq_q = w r (1 + 3 eu + 4.5 eu^2 - 1.5 u.u), S_q = w (1 - omega/2)(3 Fe + 9 eu Fe - 3 u.F). */
__global__ void k_adj_collide_srt(const float *__restrict__ f, const float *__restrict__ Fx,
                                  const float *__restrict__ Fy, const float *__restrict__ Fz,
                                  const float *__restrict__ f_l, const uint8_t *__restrict__ flags,
                                  const float *__restrict__ lam, float omega_f, float C_darcy_f,
                                  double *__restrict__ lam_f, double *__restrict__ lam_F, size_t N)
{
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N;
         c += (size_t)gridDim.x * blockDim.x) {
        double lf[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) lf[q] = 0.0;
        double lFx = 0.0, lFy = 0.0, lFz = 0.0;
        const uint8_t fg = flags[c];
        if (fg != SOLID && fg != GAS) {
            const double eps = 1e-6, om = omega_f;
            double r = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const double v = f[(size_t)q * N + c];
                r += v;
                jx += v * ex(q);
                jy += v * ey(q);
                jz += v * ez(q);
            }
            const double fx = Fx[c], fy = Fy[c], fz = Fz[c], fl = f_l[c];
            const double one_m = 1.0 - fl;
            const double ck = 1.0 + (double)C_darcy_f * (one_m * one_m) / (fl * (fl * fl) + eps);
            const double rr = r + eps, inv = 1.0 / (rr * ck);
            const double u = (jx + 0.5 * fx) * inv, v = (jy + 0.5 * fy) * inv,
                         w = (jz + 0.5 * fz) * inv;
            const double u2 = u * u + v * v + w * w;
            double lr = 0.0, lux = 0.0, luy = 0.0, luz = 0.0;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const double l = lam[(size_t)q * N + c];
                const double e0 = ex(q), e1 = ey(q), e2 = ez(q);
                const double W = q == 0 ? 1.0 / 3.0 : (q < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
                const double eu = e0 * u + e1 * v + e2 * w;
                const double Fe = e0 * fx + e1 * fy + e2 * fz;
                lf[q] = (1.0 - om) * l;
                lr += om * l * W * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * u2);
                const double a = om * l * W * r;            // feq's velocity terms
                const double b = l * W * (1.0 - 0.5 * om);  // the forcing term
                lux += a * (3.0 * e0 + 9.0 * eu * e0 - 3.0 * u) + b * (9.0 * Fe * e0 - 3.0 * fx);
                luy += a * (3.0 * e1 + 9.0 * eu * e1 - 3.0 * v) + b * (9.0 * Fe * e1 - 3.0 * fy);
                luz += a * (3.0 * e2 + 9.0 * eu * e2 - 3.0 * w) + b * (9.0 * Fe * e2 - 3.0 * fz);
                lFx += b * (3.0 * e0 + 9.0 * eu * e0 - 3.0 * u);
                lFy += b * (3.0 * e1 + 9.0 * eu * e1 - 3.0 * v);
                lFz += b * (3.0 * e2 + 9.0 * eu * e2 - 3.0 * w);
            }
            /* through u: du/dj = inv, du/dF = inv / 2, du/dr = -u / (r + eps) */
            lFx += 0.5 * lux * inv;
            lFy += 0.5 * luy * inv;
            lFz += 0.5 * luz * inv;
            lr -= (lux * u + luy * v + luz * w) / rr;
#pragma unroll
            for (int q = 0; q < Q; ++q) lf[q] += lr + (lux * ex(q) + luy * ey(q) + luz * ez(q)) * inv;
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) lam_f[(size_t)q * N + c] = lf[q];
        lam_F[c] = lFx;
        lam_F[N + c] = lFy;
        lam_F[2 * N + c] = lFz;
    }
}

/* Adjoint of stream (reference kernels.py:2991-3010), flags held: the stream copies, into each
 * population q of a non-SOLID cell c, population q of the neighbour n = c - e_q (x periodic,
 * y and z clamped), or, for a cell that is not GAS whose neighbour is SOLID, its own
 * population opp(q). Gather form: input (q, m) collects the adjoint of every output that
 * copies it: the cells c with n(c, q) = m (c = m + e_q, or m itself on a clamped face) that are
 * not SOLID and are GAS or see m not SOLID, and (opp(q), m) when m bounces back there. */
__global__ void k_adj_stream(const double *__restrict__ lam_out, double *__restrict__ lam_in,
                             const uint8_t *__restrict__ flags, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    const size_t N = d.n();
    const size_t m = d.idx(i, j, k);
    const uint8_t fm = flags[m];
    auto src_of = [&](int ci, int cj, int ck, int q, int &si, int &sj, int &sk) {
        si = ci - ex(q);
        si = si < 0 ? si + d.nx : (si >= d.nx ? si - d.nx : si);
        sj = imax(0, imin(cj - ey(q), d.ny - 1));
        sk = imax(0, imin(ck - ez(q), d.nz - 1));
    };
    for (int q = 0; q < Q; ++q) {
        double acc = 0.0;
        int ci = i + ex(q);
        ci = ci < 0 ? ci + d.nx : (ci >= d.nx ? ci - d.nx : ci);
        const int cy[2] = {j + ey(q), j}, cz[2] = {k + ez(q), k};
        const int ny_c = ey(q) != 0 ? 2 : 1, nz_c = ez(q) != 0 ? 2 : 1;
        for (int a = 0; a < ny_c; ++a) {
            const int cj = cy[a];
            if (cj < 0 || cj >= d.ny) continue;
            for (int b = 0; b < nz_c; ++b) {
                const int ck = cz[b];
                if (ck < 0 || ck >= d.nz) continue;
                int si, sj, sk;
                src_of(ci, cj, ck, q, si, sj, sk);
                if (si != i || sj != j || sk != k) continue;
                const size_t c = d.idx(ci, cj, ck);
                const uint8_t fc = flags[c];
                if (fc != SOLID && (fc == GAS || fm != SOLID)) acc += lam_out[(size_t)q * N + c];
            }
        }
        /* bounce-back at m: output opp(q) of m reads input q of m */
        const int o = opp(q);
        if (fm != SOLID && fm != GAS) {
            int si, sj, sk;
            src_of(i, j, k, o, si, sj, sk);
            if (flags[d.idx(si, sj, sk)] == SOLID) acc += lam_out[(size_t)o * N + m];
        }
        lam_in[(size_t)q * N + m] = acc;
    }
}

extern "C" {

int waam_adj_arc_sums(const uint8_t *flags, const float *f_l, const waam_adj_consts *ac,
                      double *out2, int nx, int ny, int nz, int layout, void *scratch,
                      void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_adj_arc_sums", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_adj_arc_sums", {flags, f_l, ac, out2, scratch}));
    Dims d{nx, ny, nz, layout};
    AdjC a = make_adj(*ac);
    hipStream_t s = as_stream(stream);
    double *parts = reinterpret_cast<double *>((char *)scratch + kScratchTelemetryParts);
    hipLaunchKernelGGL(k_adj_arc_partials, dim3(kAdjBlocks), dim3(kBlock), 0, s, flags, f_l, a, d,
                       parts);
    hipLaunchKernelGGL(k_adj_total, dim3(1), dim3(kBlock), 0, s, parts, kAdjBlocks, 2, out2, 0);
    return check_launch("waam_adj_arc_sums");
}

int waam_adj_terminal(const float *H, const float *grad_T, const float *grad_fl,
                      const uint8_t *flags, const waam_adj_consts *ac, float *lam, int64_t n,
                      void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_adj_terminal: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_adj_terminal", {H, grad_T, grad_fl, flags, ac, lam}));
    AdjC a = make_adj(*ac);
    hipLaunchKernelGGL(k_adj_terminal, dim3(grid_for((size_t)n)), dim3(kBlock), 0,
                       as_stream(stream), H, grad_T, grad_fl, flags, a.pc, lam, (size_t)n);
    return check_launch("waam_adj_terminal");
}

int waam_adj_thermal_step(const float *lam_next, float *lam_out, const float *H_n, const float *A,
                          const float *E, const float *F, const float *f_l_n,
                          const uint8_t *flags, const double *arc_sums,
                          const waam_adj_consts *ac, double *grads, float *lamC, int nx, int ny,
                          int nz, int layout, void *scratch, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_adj_thermal_step", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_adj_thermal_step", {lam_next, lam_out, H_n, A, E, F, f_l_n,
                                                        flags, arc_sums, ac, grads, lamC,
                                                        scratch}));
    if (lam_next == lam_out || lamC == lam_next || lamC == lam_out)
        return set_error(WAAM_ERR_INVALID, "waam_adj_thermal_step: adjoint buffers alias");
    Dims d{nx, ny, nz, layout};
    AdjC a = make_adj(*ac);
    hipStream_t s = as_stream(stream);
    size_t N = d.n();
    double *parts = reinterpret_cast<double *>((char *)scratch + kScratchTelemetryParts);
    hipLaunchKernelGGL(k_adj_step1, dim3(grid_for(N)), dim3(kBlock), 0, s, lam_next, E, F, flags,
                       a, lamC, N);
    hipLaunchKernelGGL(k_adj_step2, dim3(kAdjBlocks), dim3(kBlock), 0, s, lamC, H_n, A, f_l_n,
                       flags, arc_sums, a, d, lam_out, parts);
    hipLaunchKernelGGL(k_adj_total, dim3(1), dim3(kBlock), 0, s, parts, kAdjBlocks, 3, grads, 1);
    return check_launch("waam_adj_thermal_step");
}

int waam_adj_marangoni(const float *T, const float *phi, const float *f_l, const uint8_t *flags,
                       const float *lam_x, const float *lam_y, const float *lam_z, float dgamma,
                       double *T_adj, double *s_cell, double *G, int nx, int ny, int nz,
                       int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_adj_marangoni", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_adj_marangoni",
                              {T, phi, f_l, flags, lam_x, lam_y, lam_z, T_adj, s_cell, G}));
    Dims d{nx, ny, nz, layout};
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_adj_marangoni_project, grid3(d), block3(), 0, st, T, phi, f_l, flags,
                       lam_x, lam_y, lam_z, dgamma, G, s_cell, d);
    hipLaunchKernelGGL(k_adj_marangoni_gather, grid3(d), block3(), 0, st, (const double *)G, T_adj,
                       d);
    return check_launch("waam_adj_marangoni");
}

int waam_adj_collide_srt(const float *f_src, const float *Fx, const float *Fy, const float *Fz,
                         const float *f_l, const uint8_t *flags, const float *lam, float omega,
                         float C_darcy, double *lam_f, double *lam_F, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_adj_collide_srt: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_adj_collide_srt", {f_src, Fx, Fy, Fz, f_l, flags, lam, lam_f,
                                                       lam_F}));
    hipLaunchKernelGGL(k_adj_collide_srt, dim3(grid_for((size_t)n)), dim3(kBlock), 0,
                       as_stream(stream), f_src, Fx, Fy, Fz, f_l, flags, lam, omega, C_darcy,
                       lam_f, lam_F, (size_t)n);
    return check_launch("waam_adj_collide_srt");
}

int waam_adj_stream(const double *lam_out, double *lam_in, const uint8_t *flags, int nx, int ny,
                    int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_adj_stream", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_adj_stream", {lam_out, lam_in, flags}));
    if (lam_out == lam_in) return set_error(WAAM_ERR_INVALID, "waam_adj_stream: in place");
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_adj_stream, grid3(d), block3(), 0, as_stream(stream), lam_out, lam_in,
                       flags, d);
    return check_launch("waam_adj_stream");
}

}  // extern "C"
