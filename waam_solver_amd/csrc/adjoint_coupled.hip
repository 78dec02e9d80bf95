// adjoint_coupled.hip — binary64 reverse-mode derivative of the coupled step with frozen flags
// and phi (the differentiation form through the flow).
//
// The forward map per step (reference solvers/coupled_step.py:97-409 with the free surface, the
// flags, deposition and the stream held; with frozen flags the stream never feeds a collided
// cell, whose next collision reads its own post-collision populations):
//   A = H_n + dH_arc(eta, s_sigma); B = ceil(A)                 arc and enthalpy ceiling
//   C = B + k lap(T_n) - u_n . grad_upwind(B)                    conduction and advection
//   E = C - losses(T_n; s_h); F = max(E, floor); H_{n+1} = ceil(F)
//   T_{n+1}, f_l_{n+1} = phase(H_{n+1})
//   F_raw = const + Marangoni(T_{n+1}) + hydrostatic(f_l) + buoyancy(T, f_l); F = clamp(F_raw)
//   f_{n+1}, u_raw = MRT(f_n, F, f_l_{n+1}); u_{n+1} = mach_clamp(u_raw)
// on the FLUID/IFACE cells for the flow and every non-GAS cell for the enthalpy. The adjoint
// state (lambda_H, lambda_f, lambda_u) is binary64 throughout; derivatives at piecewise points
// are one-sided on the branch the forward took (clamps give 0, the minmod limiter the slope of
// the difference it selected, the phase map the slope of its side; at u = 0 the upwind side
// the forward used).
#include "arc_cell.h"
#include "thermal_cell.h"

namespace waam {

/* binary64 value with one forward-mode tangent */
struct D1 {
    double v, d;
};
__device__ __forceinline__ D1 dk(double v) { return D1{v, 0.0}; }
__device__ __forceinline__ D1 operator+(D1 a, D1 b) { return D1{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ D1 operator-(D1 a, D1 b) { return D1{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ D1 operator-(D1 a) { return D1{-a.v, -a.d}; }
__device__ __forceinline__ D1 operator*(D1 a, D1 b) { return D1{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ D1 operator/(D1 a, D1 b)
{
    const double q = a.v / b.v;
    return D1{q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ D1 operator*(double s, D1 a) { return D1{s * a.v, s * a.d}; }
__device__ __forceinline__ D1 operator+(double s, D1 a) { return D1{s + a.v, a.d}; }

/* The central-moment MRT map of lbm_cell.h (reference cumulant_kernel.py:59-209) in binary64 on
 * the ten moments it reads of the populations — m = (r, jx, jy, jz, Pxx, Pyy, Pzz, Pxy, Pxz,
 * Pyz) — the force and f_l: out[q] the post-collision populations, u the velocity before the
 * Mach clamp. The reconstruction weights are the reference's truncated decimals. */
__device__ __forceinline__ void mrt_map(const D1 (&m)[10], D1 fx, D1 fy, D1 fz, D1 fl,
                                        double omega_s, double omega_b, double C_darcy,
                                        D1 (&out)[Q], D1 (&u3)[3])
{
    const double cs2 = 1.0 / 3.0, inv_cs2 = 3.0, eps = 1.0e-6;
    const D1 r = m[0];
    const D1 inv_rho = dk(1.0) / (r + dk(eps));
    const D1 uxr = (m[1] + 0.5 * fx) * inv_rho;
    const D1 uyr = (m[2] + 0.5 * fy) * inv_rho;
    const D1 uzr = (m[3] + 0.5 * fz) * inv_rho;
    const D1 one_m = dk(1.0) - fl;
    const D1 ck = 1.0 + C_darcy * (one_m * one_m) / (fl * (fl * fl) + dk(eps));
    const D1 u = uxr / ck, v = uyr / ck, w = uzr / ck;
    const D1 kxx = m[4] - r * u * u, kyy = m[5] - r * v * v, kzz = m[6] - r * w * w;
    const D1 kxy = m[7] - r * u * v, kxz = m[8] - r * u * w, kyz = m[9] - r * v * w;
    const D1 kmean = (1.0 / 3.0) * (kxx + kyy + kzz);
    const D1 kmean_p = kmean - omega_b * (kmean - cs2 * r);
    const double sk = 1.0 - omega_s;
    const D1 kxxp = kmean_p + sk * (kxx - kmean), kyyp = kmean_p + sk * (kyy - kmean),
             kzzp = kmean_p + sk * (kzz - kmean);
    const D1 kxyp = sk * kxy, kxzp = sk * kxz, kyzp = sk * kyz;
    const D1 u2 = u * u + v * v + w * w;
    const D1 uF = u * fx + v * fy + w * fz;
    const D1 rcs = cs2 * r;
    const D1 dxx = kxxp - rcs, dyy = kyyp - rcs, dzz = kzzp - rcs;
    const double s_pref = 1.0 - 0.5 * omega_s;
    const double k45 = 0.5 * inv_cs2 * inv_cs2;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double e0 = ex(q), e1 = ey(q), e2 = ez(q);
        const D1 eu = e0 * u + e1 * v + e2 * w;
        const D1 eF = e0 * fx + e1 * fy + e2 * fz;
        D1 K;
        switch (q) {
            case 0: K = dk(0.0); break;
            case 1: case 2: K = dxx; break;
            case 3: case 4: K = dyy; break;
            case 5: case 6: K = dzz; break;
            case 7: case 8: K = (dxx + dyy) + 2.0 * kxyp; break;
            case 9: case 10: K = (dxx + dyy) - 2.0 * kxyp; break;
            case 11: case 12: K = (dxx + dzz) + 2.0 * kxzp; break;
            case 13: case 14: K = (dxx + dzz) - 2.0 * kxzp; break;
            case 15: case 16: K = (dyy + dzz) + 2.0 * kyzp; break;
            default: K = (dyy + dzz) - 2.0 * kyzp; break;
        }
        const double wq_ = q == 0 ? 0.3333333333 : (q < 7 ? 0.0555555556 : 0.0277777778);
        const D1 feq = wq_ * (r + inv_cs2 * r * eu + (0.5 * inv_cs2 * inv_cs2) * r * eu * eu -
                              (0.5 * inv_cs2) * r * u2 + k45 * K);
        const D1 S = (wq_ * s_pref) * (inv_cs2 * eF + (inv_cs2 * inv_cs2) * eF * eu - inv_cs2 * uF);
        out[q] = feq + S;
    }
    u3[0] = u;
    u3[1] = v;
    u3[2] = w;
}

/* d(moment t)/d(population q): r, j, the second moments (0 beyond the ten moments) */
__device__ __forceinline__ double moment_row(int t, int q)
{
    const double e0 = ex(q), e1 = ey(q), e2 = ez(q);
    switch (t) {
        case 0: return 1.0;
        case 1: return e0;
        case 2: return e1;
        case 3: return e2;
        case 4: return e0 * e0;
        case 5: return e1 * e1;
        case 6: return e2 * e2;
        case 7: return e0 * e1;
        case 8: return e0 * e2;
        default: return e1 * e2;
    }
}

/* Jacobian-transpose product of v_out = v * cap / (|v| + 1e-12) where |v| > cap, identity else */
__device__ __forceinline__ void clamp_adj(double vx, double vy, double vz, double cap,
                                          double &lx, double &ly, double &lz)
{
    if (!(cap > 0.0)) return;
    const double mag = sqrt(vx * vx + vy * vy + vz * vz);
    if (!(mag > cap)) return;
    const double den = mag + 1e-12, s = cap / den;
    /* ds/dv_j = -cap / den^2 * v_j / mag */
    const double lv = lx * vx + ly * vy + lz * vz;
    const double k = -cap / (den * den) / mag * lv;
    lx = s * lx + k * vx;
    ly = s * ly + k * vy;
    lz = s * lz + k * vz;
}

/* Collision adjoint on FLUID/IFACE cells (zero elsewhere): for lam_f (post-collision
 * populations) and lam_u (velocity after the Mach clamp), lam_fin = dL/d(pre-collision
 * populations), lam_F = dL/d(force the collision saw), lam_fl += dL/d(f_l) through the drag.
 * Forward-mode over the 14 inputs of the map (ten moments, force, f_l), one tangent at a time. */
__global__ void k_cadj_collide_mrt(const float *__restrict__ f, const float *__restrict__ Fx,
                                   const float *__restrict__ Fy, const float *__restrict__ Fz,
                                   const float *__restrict__ f_l, const uint8_t *__restrict__ flags,
                                   const double *__restrict__ lam_f, const double *__restrict__ lam_u,
                                   double omega_s, double omega_b, double C_darcy, double u_cap,
                                   double *__restrict__ lam_fin, double *__restrict__ lam_F,
                                   double *__restrict__ lam_fl, size_t N)
{
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N;
         c += (size_t)gridDim.x * blockDim.x) {
        const uint8_t fg = flags[c];
        if (fg == SOLID || fg == GAS) {
#pragma unroll
            for (int q = 0; q < Q; ++q) lam_fin[(size_t)q * N + c] = 0.0;
            lam_F[c] = lam_F[N + c] = lam_F[2 * N + c] = 0.0;
            continue;
        }
        double mv[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int q = 0; q < Q; ++q) {
            const double x = f[(size_t)q * N + c];
#pragma unroll
            for (int t = 0; t < 10; ++t) mv[t] += moment_row(t, q) * x;
        }
        const double fx = Fx[c], fy = Fy[c], fz = Fz[c], fl = f_l[c];
        /* the velocity adjoint through the Mach clamp, at the map's unclamped velocity */
        D1 m0[10], out[Q], u3[3];
#pragma unroll
        for (int t = 0; t < 10; ++t) m0[t] = dk(mv[t]);
        mrt_map(m0, dk(fx), dk(fy), dk(fz), dk(fl), omega_s, omega_b, C_darcy, out, u3);
        double lu[3] = {lam_u[c], lam_u[N + c], lam_u[2 * N + c]};
        clamp_adj(u3[0].v, u3[1].v, u3[2].v, u_cap, lu[0], lu[1], lu[2]);
        double lf[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) lf[q] = lam_f[(size_t)q * N + c];
        double g[14];
        for (int t = 0; t < 14; ++t) {
            D1 m[10];
#pragma unroll
            for (int s = 0; s < 10; ++s) m[s] = D1{mv[s], s == t ? 1.0 : 0.0};
            const D1 dfx{fx, t == 10 ? 1.0 : 0.0}, dfy{fy, t == 11 ? 1.0 : 0.0},
                dfz{fz, t == 12 ? 1.0 : 0.0}, dfl{fl, t == 13 ? 1.0 : 0.0};
            mrt_map(m, dfx, dfy, dfz, dfl, omega_s, omega_b, C_darcy, out, u3);
            double acc = lu[0] * u3[0].d + lu[1] * u3[1].d + lu[2] * u3[2].d;
#pragma unroll
            for (int q = 0; q < Q; ++q) acc += lf[q] * out[q].d;
            g[t] = acc;
        }
        for (int q = 0; q < Q; ++q) {
            double a = 0.0;
#pragma unroll
            for (int t = 0; t < 10; ++t) a += g[t] * moment_row(t, q);
            lam_fin[(size_t)q * N + c] = a;
        }
        lam_F[c] = g[10];
        lam_F[N + c] = g[11];
        lam_F[2 * N + c] = g[12];
        lam_fl[c] += g[13];
    }
}

/* Force adjoint on FLUID/IFACE cells: through the magnitude clamp at the raw force, then the
 * hydrostatic and buoyancy terms (reference kernels.py:2232-2268 order: both apply where
 * f_l > 0): lam_Fraw (3 arrays, for the Marangoni adjoint), lam_T += and lam_fl +=. */
__global__ void k_cadj_forces(const float *__restrict__ Frx, const float *__restrict__ Fry,
                              const float *__restrict__ Frz, const float *__restrict__ T,
                              const float *__restrict__ f_l, const uint8_t *__restrict__ flags,
                              const double *__restrict__ lam_F, double F_cap, double g_lu,
                              double beta, double T_l, int enable_hydro,
                              double *__restrict__ lam_Fraw, double *__restrict__ lam_T,
                              double *__restrict__ lam_fl, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const uint8_t fg = flags[c];
    double lx = 0.0, ly = 0.0, lz = 0.0;
    if (fg != SOLID && fg != GAS) {
        lx = lam_F[c];
        ly = lam_F[N + c];
        lz = lam_F[2 * N + c];
        clamp_adj(Frx[c], Fry[c], Frz[c], F_cap, lx, ly, lz);
        const double fl = f_l[c];
        if (fl > 0.0) {
            double dfl = g_lu * beta * ((double)T[c] - T_l);
            if (enable_hydro) dfl -= g_lu;
            lam_fl[c] += lz * dfl;
            lam_T[c] += lz * g_lu * beta * fl;
        }
    }
    lam_Fraw[c] = lx;
    lam_Fraw[N + c] = ly;
    lam_Fraw[2 * N + c] = lz;
}

/* Marangoni adjoint, projection pass with binary64 lambda (see adjoint.hip) */
__global__ void k_cadj_marangoni_project(const float *__restrict__ phi,
                                         const float *__restrict__ f_l,
                                         const uint8_t *__restrict__ flags,
                                         const double *__restrict__ lam, double dgamma,
                                         double *__restrict__ G, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    const size_t N = d.n();
    const size_t c = d.idx(i, j, k);
    double gx = 0.0, gy = 0.0, gz = 0.0;
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
            const double ax = lam[c], ay = lam[N + c], az = lam[2 * N + c];
            const double nl = nx * ax + ny * ay + nz * az;
            const double w = dgamma * (double)gm;
            gx = w * (ax - nl * nx);
            gy = w * (ay - nl * ny);
            gz = w * (az - nl * nz);
        }
    }
    G[c] = gx;
    G[N + c] = gy;
    G[2 * N + c] = gz;
}

/* gather of the projected Marangoni adjoint over the clamped central differences (as in
 * adjoint.hip): T_adj += */
__global__ void k_cadj_marangoni_gather(const double *__restrict__ G, double *__restrict__ T_adj,
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

/* lam_H += lam_T dT/dH + lam_fl df_l/dH at H (phase slopes of the side H is on); 0 on GAS */
__global__ void k_cadj_phase(const float *__restrict__ H, const uint8_t *__restrict__ flags,
                             PhaseConsts pc, const double *__restrict__ lam_T,
                             const double *__restrict__ lam_fl, double *__restrict__ lam_H,
                             size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    if (flags[c] == GAS) {
        lam_H[c] = 0.0;
        return;
    }
    const float h = H[c];
    double dT, dfl;
    if (h <= pc.H_sol || h >= pc.H_liq) {
        dT = 1.0 / (double)pc.cp_rho;
        dfl = 0.0;
    } else {
        dfl = 1.0 / (double)pc.L_rho;
        dT = ((double)pc.T_l - (double)pc.T_s) / (double)pc.L_rho;
    }
    lam_H[c] += lam_T[c] * dT + lam_fl[c] * dfl;
}

struct CAdj {
    PhaseConsts pc;
    double cap_l, cap_s, h_floor, kcond, loss_geo, s_h;
    LossConsts lc0;
    ArcParams ap;
    double sigma, sigma0, K, eta;
    int enable_cap, enable_loss, enable_arc;
};

static CAdj make_cadj(const waam_adj_consts &a)
{
    CAdj c{};
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
    c.K = (double)(a.Q_w * a.eta * a.dt / a.dx3);
    c.eta = a.eta;
    c.enable_cap = a.enable_cap;
    c.enable_loss = a.enable_loss;
    c.enable_arc = a.enable_arc;
    return c;
}

/* dq/dT of the loss flux (q is evaluated at min(T, 2500)) */
__device__ __forceinline__ float loss_slope_f(float T, const LossConsts &lc)
{
    if (!(T < 2500.0f)) return 0.0f;
    float s = 0.0f;
    if (lc.enable_conv == 1) s += lc.h_conv;
    if (lc.enable_rad == 1) s += 4.0f * lc.eps_rad * lc.sigma_sb * (T * T * T);
    return s;
}

/* the forward's ceiling clamps h (the branch the forward took) */
__device__ __forceinline__ bool cceil(float h, const CAdj &a)
{
    const float cap_l = (float)a.cap_l, cap_s = (float)a.cap_s;
    return h > a.pc.H_liq ? h > cap_l : h > cap_s;
}

/* thermal adjoint, part 1: through the second ceiling and the floor; lamC = dL/dC and
 * lamB = lamC (the identity part of C's dependence on B); 0 on GAS */
__global__ void k_cadj_step1(const double *__restrict__ lam_next, const float *__restrict__ E,
                             const float *__restrict__ F, const uint8_t *__restrict__ flags, CAdj a,
                             double *__restrict__ lamC, double *__restrict__ lamB, size_t N)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    double l = 0.0;
    if (flags[c] != GAS) {
        l = lam_next[c];
        if (a.enable_cap && cceil(F[c], a)) l = 0.0;
        if (a.enable_loss && E[c] < (float)a.h_floor) l = 0.0;
    }
    lamC[c] = l;
    lamB[c] = l;
}

/* the masked enthalpy operand of the upwind stencil: value and source cell (clamped index; a
 * GAS cell gives the fallback's value and source) */
__device__ __forceinline__ void masked_src(const float *__restrict__ B,
                                           const uint8_t *__restrict__ flags, int i, int j, int k,
                                           float fb_v, size_t fb_c, const Dims &d, float &v,
                                           size_t &src)
{
    const int ii = imax(0, imin(i, d.nx - 1));
    const int jj = imax(0, imin(j, d.ny - 1));
    const int kk = imax(0, imin(k, d.nz - 1));
    const size_t c = d.idx(ii, jj, kk);
    if (flags[c] == GAS) {
        v = fb_v;
        src = fb_c;
    } else {
        v = B[c];
        src = c;
    }
}

/* d(upwind_grad)/d(qmm, qm, qc, qp, qpp) on the branch the forward took */
__device__ __forceinline__ void upwind_grad_slopes(float u, float qmm, float qm, float qc,
                                                   float qp, float qpp, double (&s)[5])
{
    s[0] = s[1] = s[2] = s[3] = s[4] = 0.0;
    if (u > 0.0f) {
        const float d1 = qc - qm;
        const float d2 = 0.5f * (3.0f * qc - 4.0f * qm + qmm);
        if (d1 * d2 > 0.0f) {
            if (fabsf(d1) < fabsf(d2)) {
                s[2] = 1.0;
                s[1] = -1.0;
            } else {
                s[2] = 1.5;
                s[1] = -2.0;
                s[0] = 0.5;
            }
        }
    } else {
        const float d1 = qp - qc;
        const float d2 = 0.5f * (-3.0f * qc + 4.0f * qp - qpp);
        if (d1 * d2 > 0.0f) {
            if (fabsf(d1) < fabsf(d2)) {
                s[3] = 1.0;
                s[2] = -1.0;
            } else {
                s[2] = -1.5;
                s[3] = 2.0;
                s[4] = -0.5;
            }
        }
    }
}

/* thermal adjoint, part 2: the advection term of the non-SOLID cells (C -= u . g(B)): lam_u =
 * -g lamC, and lamB of every operand's source cell += -u dg/dB lamC (scattered) */
__global__ void k_cadj_advect(const double *__restrict__ lamC, const float *__restrict__ B,
                              const float *__restrict__ ux, const float *__restrict__ uy,
                              const float *__restrict__ uz, const uint8_t *__restrict__ flags,
                              double *__restrict__ lamB, double *__restrict__ lam_u, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    const size_t N = d.n();
    const size_t c = d.idx(i, j, k);
    const uint8_t fg = flags[c];
    double lu[3] = {0.0, 0.0, 0.0};
    const double l = lamC[c];
    if (fg != SOLID && fg != GAS && l != 0.0) {
        const float uc[3] = {ux[c], uy[c], uz[c]};
        const float Bc = B[c];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int o[3] = {a == 0, a == 1, a == 2};
            float vm, vp, vmm, vpp;
            size_t sm, sp, smm, spp;
            masked_src(B, flags, i - o[0], j - o[1], k - o[2], Bc, c, d, vm, sm);
            masked_src(B, flags, i + o[0], j + o[1], k + o[2], Bc, c, d, vp, sp);
            masked_src(B, flags, i - 2 * o[0], j - 2 * o[1], k - 2 * o[2], vm, sm, d, vmm, smm);
            masked_src(B, flags, i + 2 * o[0], j + 2 * o[1], k + 2 * o[2], vp, sp, d, vpp, spp);
            const float gv = upwind_grad(uc[a], vmm, vm, Bc, vp, vpp);
            lu[a] = -(double)gv * l;
            double s[5];
            upwind_grad_slopes(uc[a], vmm, vm, Bc, vp, vpp, s);
            const double f = -(double)uc[a] * l;
            const size_t src[5] = {smm, sm, c, sp, spp};
#pragma unroll
            for (int t = 0; t < 5; ++t)
                if (s[t] != 0.0) atomicAdd(&lamB[src[t]], f * s[t]);
        }
    }
    lam_u[c] = lu[0];
    lam_u[N + c] = lu[1];
    lam_u[2 * N + c] = lu[2];
}

/* thermal adjoint, part 3: conduction (symmetric operator on lamC), losses (on lamC), the first
 * ceiling and the arc (on lamB); lam_out = dL/dH_n; block partial sums of dL/d(eta, s_h,
 * s_sigma) */
__global__ void __launch_bounds__(kBlock) k_cadj_step2(
    const double *__restrict__ lamC, const double *__restrict__ lamB, const float *__restrict__ H_n,
    const float *__restrict__ A, const float *__restrict__ fl_n, const uint8_t *__restrict__ flags,
    const double *__restrict__ arc_sums, CAdj a, Dims d, double *__restrict__ lam_out,
    double *__restrict__ parts)
{
    __shared__ double sh_s[kBlock], sh_c[kBlock];
    double g[3] = {0.0, 0.0, 0.0}, gc[3] = {0.0, 0.0, 0.0};
    const size_t N = d.n();
    const double W = a.enable_arc ? arc_sums[0] : 0.0;
    const double S = a.enable_arc ? arc_sums[1] : 0.0;
    const double epw = (a.enable_arc && (float)W > 1e-9f) ? a.K / W : 0.0;
    const double rbar = W > 0.0 ? S / W : 0.0;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N;
         c += (size_t)gridDim.x * blockDim.x) {
        if (flags[c] == GAS) {
            lam_out[c] = 0.0;
            continue;
        }
        int i, j, k;
        d.ijk(c, i, j, k);
        const double lc = lamC[c];
        const float h = H_n[c];
        float Tn, fln;
        phase_of(h, a.pc, Tn, fln);
        double acc = 0.0;
        const int di[6] = {-1, 1, 0, 0, 0, 0}, dj[6] = {0, 0, -1, 1, 0, 0},
                  dkk[6] = {0, 0, 0, 0, -1, 1};
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            int ii = i + di[t], jj = j + dj[t], kk = k + dkk[t];
            if (ii < 0 || ii >= d.nx || jj < 0 || jj >= d.ny || kk < 0 || kk >= d.nz) continue;
            size_t n = d.idx(ii, jj, kk);
            if (flags[n] == GAS) continue;
            acc += lamC[n] - lc;
        }
        double lT = a.kcond * acc;
        if (a.enable_loss) {
            int n_exp = exposed_faces(flags, i, j, k, d);
            if (n_exp > 0) {
                LossConsts ls = a.lc0;
                ls.h_conv = a.lc0.h_conv * (float)a.s_h;
                ls.eps_rad = a.lc0.eps_rad * (float)a.s_h;
                const double geo = (double)n_exp * a.loss_geo;
                lT -= lc * geo * (double)loss_slope_f(Tn, ls);
                double x = -lc * geo * (double)loss_flux(Tn, a.lc0), ss, e;
                two_sum(g[1], x, ss, e);
                g[1] = ss;
                gc[1] += e;
            }
        }
        const double lA = (a.enable_cap && cceil(A[c], a)) ? 0.0 : lamB[c];
        double dT;
        if (h <= a.pc.H_sol || h >= a.pc.H_liq) dT = 1.0 / (double)a.pc.cp_rho;
        else dT = ((double)a.pc.T_l - (double)a.pc.T_s) / (double)a.pc.L_rho;
        lam_out[c] = lA + lT * dT;
        if (a.enable_arc && epw > 0.0 && lA != 0.0) {
            float wg, wdep;
            if (arc_cell_weights(a.ap, i, j, k, fl_n[c], wg, wdep)) {
                float fi = (float)i - a.ap.arc_i, fj = (float)j - a.ap.arc_j,
                      fk = (float)k - a.ap.arc_k;
                float r2 = fi * fi + fj * fj + fk * fk * 0.25f;
                const double dh = epw * (double)wg * (double)wdep;
                double ss, e;
                two_sum(g[0], lA * dh / a.eta, ss, e);
                g[0] = ss;
                gc[0] += e;
                const double s3 = a.sigma * a.sigma * a.sigma;
                two_sum(g[2], lA * dh * a.sigma0 * ((double)r2 - rbar) / s3, ss, e);
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

/* fixed-order total of kAdjBlocks groups of 3 compensated pairs, accumulated into out[0..2] */
__global__ void __launch_bounds__(kBlock) k_cadj_total(const double *__restrict__ parts, int npart,
                                                       double *__restrict__ out)
{
    __shared__ double sh_s[kBlock], sh_c[kBlock];
    for (int g = 0; g < 3; ++g) {
        double s = 0.0, cmp = 0.0;
        for (int b = threadIdx.x; b < npart; b += blockDim.x) {
            double ss, e;
            two_sum(s, parts[(size_t)b * 6 + 2 * g], ss, e);
            s = ss;
            cmp += parts[(size_t)b * 6 + 2 * g + 1] + e;
        }
        block_reduce_pair(s, cmp, sh_s, sh_c);
        if (threadIdx.x == 0) out[g] += s + cmp;
        __syncthreads();
    }
}

constexpr int kCAdjBlocks = 1024;

}  // namespace waam

using namespace waam;

extern "C" {

int waam_cadj_collide_mrt(const float *f_src, const float *Fx, const float *Fy, const float *Fz,
                          const float *f_l, const uint8_t *flags, const double *lam_f,
                          const double *lam_u, double omega_s, double omega_b, double C_darcy,
                          double u_cap, double *lam_fin, double *lam_F, double *lam_fl, int64_t n,
                          void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_cadj_collide_mrt: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_cadj_collide_mrt", {f_src, Fx, Fy, Fz, f_l, flags, lam_f,
                                                        lam_u, lam_fin, lam_F, lam_fl}));
    if (lam_fin == lam_f)
        return set_error(WAAM_ERR_INVALID, "waam_cadj_collide_mrt: lam_fin aliases lam_f");
    unsigned g = grid_for((size_t)n);
    if (g > 4096u) g = 4096u;
    hipLaunchKernelGGL(k_cadj_collide_mrt, dim3(g), dim3(kBlock), 0, as_stream(stream), f_src,
                       Fx, Fy, Fz, f_l, flags, lam_f, lam_u, omega_s, omega_b, C_darcy, u_cap,
                       lam_fin, lam_F, lam_fl, (size_t)n);
    return check_launch("waam_cadj_collide_mrt");
}

int waam_cadj_forces(const float *Frx, const float *Fry, const float *Frz, const float *T,
                     const float *f_l, const uint8_t *flags, const double *lam_F, double F_cap,
                     double g_lu, double beta, double T_l, int enable_hydro, double *lam_Fraw,
                     double *lam_T, double *lam_fl, int64_t n, void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_cadj_forces: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_cadj_forces",
                              {Frx, Fry, Frz, T, f_l, flags, lam_F, lam_Fraw, lam_T, lam_fl}));
    hipLaunchKernelGGL(k_cadj_forces, dim3(grid_for((size_t)n)), dim3(kBlock), 0,
                       as_stream(stream), Frx, Fry, Frz, T, f_l, flags, lam_F, F_cap, g_lu, beta,
                       T_l, enable_hydro, lam_Fraw, lam_T, lam_fl, (size_t)n);
    return check_launch("waam_cadj_forces");
}

int waam_cadj_marangoni(const float *phi, const float *f_l, const uint8_t *flags,
                        const double *lam, double dgamma, double *G, double *T_adj, int nx, int ny,
                        int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_cadj_marangoni", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_cadj_marangoni", {phi, f_l, flags, lam, G, T_adj}));
    Dims d{nx, ny, nz, layout};
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_cadj_marangoni_project, grid3(d), block3(), 0, st, phi, f_l, flags, lam,
                       dgamma, G, d);
    hipLaunchKernelGGL(k_cadj_marangoni_gather, grid3(d), block3(), 0, st, (const double *)G,
                       T_adj, d);
    return check_launch("waam_cadj_marangoni");
}

int waam_cadj_phase(const float *H, const uint8_t *flags, const waam_adj_consts *ac,
                    const double *lam_T, const double *lam_fl, double *lam_H, int64_t n,
                    void *stream)
{
    if (n <= 0) return set_error(WAAM_ERR_INVALID, "waam_cadj_phase: n <= 0");
    WAAM_RETURN_IF(check_ptrs("waam_cadj_phase", {H, flags, ac, lam_T, lam_fl, lam_H}));
    CAdj a = make_cadj(*ac);
    hipLaunchKernelGGL(k_cadj_phase, dim3(grid_for((size_t)n)), dim3(kBlock), 0,
                       as_stream(stream), H, flags, a.pc, lam_T, lam_fl, lam_H, (size_t)n);
    return check_launch("waam_cadj_phase");
}

int waam_cadj_thermal_step(const double *lam_next, double *lam_out, double *lam_u,
                           const float *H_n, const float *A, const float *B, const float *E,
                           const float *F, const float *f_l_n, const float *ux, const float *uy,
                           const float *uz, const uint8_t *flags, const double *arc_sums,
                           const waam_adj_consts *ac, double *grads, double *lamC, double *lamB,
                           int nx, int ny, int nz, int layout, void *scratch, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_cadj_thermal_step", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_cadj_thermal_step",
                              {lam_next, lam_out, lam_u, H_n, A, B, E, F, f_l_n, ux, uy, uz, flags,
                               arc_sums, ac, grads, lamC, lamB, scratch}));
    if (lam_next == lam_out || lamC == lamB || lamC == lam_next || lamB == lam_next ||
        lamC == lam_out || lamB == lam_out)
        return set_error(WAAM_ERR_INVALID, "waam_cadj_thermal_step: adjoint buffers alias");
    Dims d{nx, ny, nz, layout};
    CAdj a = make_cadj(*ac);
    hipStream_t s = as_stream(stream);
    size_t N = d.n();
    double *parts = reinterpret_cast<double *>((char *)scratch + kScratchTelemetryParts);
    hipLaunchKernelGGL(k_cadj_step1, dim3(grid_for(N)), dim3(kBlock), 0, s, lam_next, E, F, flags,
                       a, lamC, lamB, N);
    hipLaunchKernelGGL(k_cadj_advect, grid3(d), block3(), 0, s, (const double *)lamC, B, ux, uy,
                       uz, flags, lamB, lam_u, d);
    hipLaunchKernelGGL(k_cadj_step2, dim3(kCAdjBlocks), dim3(kBlock), 0, s, (const double *)lamC,
                       (const double *)lamB, H_n, A, f_l_n, flags, arc_sums, a, d, lam_out, parts);
    hipLaunchKernelGGL(k_cadj_total, dim3(1), dim3(kBlock), 0, s, (const double *)parts,
                       kCAdjBlocks, grads);
    return check_launch("waam_cadj_thermal_step");
}

}  // extern "C"
