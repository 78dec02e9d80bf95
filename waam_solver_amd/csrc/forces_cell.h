/*
 * forces_cell.h — per-cell body-force terms shared by the standalone force kernels and the
 * fused collide kernel. Each returns the increment the reference adds with += so a caller can
 * accumulate terms in the reference's order (coupled_step.py:298-362).
 */
#pragma once

#include "waam_common.h"

namespace waam {

/* Loads first, arithmetic after: every routine below gathers all the values it may need with
 * clamped indices before it evaluates anything, and selects between them instead of branching
 * around loads, so a cell's loads issue together rather than as a dependent chain. */

/* index offsets of the six face neighbours of (i, j, k) with the reference's clamps, in the
 * order x+, x-, y+, y-, z+, z- */
__device__ __forceinline__ void face_idx(int i, int j, int k, const Dims &d, size_t (&o)[6])
{
    o[0] = d.idx(imin(i + 1, d.nx - 1), j, k);
    o[1] = d.idx(imax(i - 1, 0), j, k);
    o[2] = d.idx(i, imin(j + 1, d.ny - 1), k);
    o[3] = d.idx(i, imax(j - 1, 0), k);
    o[4] = d.idx(i, j, imin(k + 1, d.nz - 1));
    o[5] = d.idx(i, j, imax(k - 1, 0));
}

/* _phi_unit_normal_at (reference kernels.py:769-796) from the six loaded face values */
__device__ __forceinline__ void normal_from(const float (&v)[6], float eps, float &nx_, float &ny_,
                                            float &nz_, float &g)
{
    float dpx = 0.5f * (v[0] - v[1]);
    float dpy = 0.5f * (v[2] - v[3]);
    float dpz = 0.5f * (v[4] - v[5]);
    g = sqrtf(dpx * dpx + dpy * dpy + dpz * dpz);
    bool ok = g >= eps;
    nx_ = ok ? dpx / g : 0.0f;
    ny_ = ok ? dpy / g : 0.0f;
    nz_ = ok ? dpz / g : 0.0f;
}

/* phi on the stencil of the Brackbill curvature: the face values of the cell and of each of
 * its six (clamped) face neighbours */
struct PhiStencil {
    float c[6];      // faces of the cell
    float n[6][6];   // faces of face neighbour m (m in face_idx order)
};

__device__ __forceinline__ void load_phi_stencil(const float *__restrict__ phi, int i, int j,
                                                 int k, const Dims &d, PhiStencil &st)
{
    size_t o[6];
    face_idx(i, j, k, d, o);
    const int ni[6] = {imin(i + 1, d.nx - 1), imax(i - 1, 0), i, i, i, i};
    const int nj[6] = {j, j, imin(j + 1, d.ny - 1), imax(j - 1, 0), j, j};
    const int nk[6] = {k, k, k, k, imin(k + 1, d.nz - 1), imax(k - 1, 0)};
    size_t on[6][6];
#pragma unroll
    for (int m = 0; m < 6; ++m) face_idx(ni[m], nj[m], nk[m], d, on[m]);
#pragma unroll
    for (int q = 0; q < 6; ++q) st.c[q] = phi[o[q]];
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
        for (int q = 0; q < 6; ++q) st.n[m][q] = phi[on[m][q]];
}

/* _brackbill_curvature_at — reference kernels.py:799-855 */
__device__ __forceinline__ void brackbill_kappa(const PhiStencil &st, float eps, float &kappa,
                                                float &g0)
{
    float n0x, n0y, n0z;
    normal_from(st.c, eps, n0x, n0y, n0z, g0);
    float a, b, nxp, nxm, nyp, nym, nzp, nzm, gp, gm, gp2, gm2, gp3, gm3;
    normal_from(st.n[0], eps, nxp, a, b, gp);
    normal_from(st.n[1], eps, nxm, a, b, gm);
    normal_from(st.n[2], eps, a, nyp, b, gp2);
    normal_from(st.n[3], eps, a, nym, b, gm2);
    normal_from(st.n[4], eps, a, b, nzp, gp3);
    normal_from(st.n[5], eps, a, b, nzm, gm3);
    nxp = gp < eps ? n0x : nxp;
    nxm = gm < eps ? n0x : nxm;
    nyp = gp2 < eps ? n0y : nyp;
    nym = gm2 < eps ? n0y : nym;
    nzp = gp3 < eps ? n0z : nzp;
    nzm = gm3 < eps ? n0z : nzm;
    float div_n = 0.5f * (nxp - nxm) + 0.5f * (nyp - nym) + 0.5f * (nzp - nzm);
    kappa = g0 >= eps ? -div_n : 0.0f;
}

/* _solid_wall_normal — reference kernels.py:680-719, from the six loaded face flags */
__device__ __forceinline__ bool wall_normal(const uint8_t (&fn)[6], int i, int j, int k,
                                            const Dims &d, float &wx, float &wy, float &wz)
{
    bool sxm = i > 0 && fn[1] == SOLID, sxp = i < d.nx - 1 && fn[0] == SOLID;
    bool sym = j > 0 && fn[3] == SOLID, syp = j < d.ny - 1 && fn[2] == SOLID;
    bool szm = k > 0 && fn[5] == SOLID, szp = k < d.nz - 1 && fn[4] == SOLID;
    wx = 0.0f;
    wy = 0.0f;
    wz = 0.0f;
    if (sxm) wx += 1.0f;
    if (sxp) wx -= 1.0f;
    if (sym) wy += 1.0f;
    if (syp) wy -= 1.0f;
    if (szm) wz += 1.0f;
    if (szp) wz -= 1.0f;
    float m = sqrtf(wx * wx + wy * wy + wz * wz);
    if (m > 1e-8f) {
        wx /= m;
        wy /= m;
        wz /= m;
    }
    return sxm || sxp || sym || syp || szm || szp;
}

/* _correct_normal_contact_angle — reference kernels.py:722-766 */
__device__ __forceinline__ void contact_normal(float nxn, float nyn, float nzn, float wx, float wy,
                                               float wz, float cos_t, float sin_t, float &cx,
                                               float &cy, float &cz)
{
    float dot = nxn * wx + nyn * wy + nzn * wz;
    float tx = nxn - dot * wx, ty = nyn - dot * wy, tz = nzn - dot * wz;
    float tmag = sqrtf(tx * tx + ty * ty + tz * tz);
    if (tmag > 1e-8f) {
        cx = cos_t * wx + sin_t * tx / tmag;
        cy = cos_t * wy + sin_t * ty / tmag;
        cz = cos_t * wz + sin_t * tz / tmag;
    } else if (fabsf(wz) > 0.5f) {
        cx = sin_t;
        cy = 0.0f;
        cz = cos_t;
    } else if (fabsf(wx) > 0.5f) {
        cx = cos_t;
        cy = sin_t;
        cz = 0.0f;
    } else {
        cx = cos_t * wx + sin_t;
        cy = cos_t * wy;
        cz = cos_t * wz;
    }
    float m = sqrtf(cx * cx + cy * cy + cz * cz);
    if (m > 1e-8f) {
        cx /= m;
        cy /= m;
        cz /= m;
    }
}

/* Everything the body forces of one cell read around it: the curvature stencil of phi, and T
 * and the flags on the six faces. */
struct ForceStencil {
    PhiStencil phi;
    float T[6];
    uint8_t flags[6];
};

__device__ __forceinline__ void load_force_stencil(const float *__restrict__ phi,
                                                   const float *__restrict__ T,
                                                   const uint8_t *__restrict__ flags, int i, int j,
                                                   int k, const Dims &d, ForceStencil &st)
{
    load_phi_stencil(phi, i, j, k, d, st.phi);
    size_t o[6];
    face_idx(i, j, k, d, o);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        st.T[q] = T[o[q]];
        st.flags[q] = flags[o[q]];
    }
}

/* compute_csf_tension body — reference kernels.py:885-926. Returns false when the reference
 * adds nothing (rim, |grad phi| < eps). Caller has checked the cell is neither SOLID nor GAS. */
__device__ __forceinline__ bool csf_force(const ForceStencil &fs, int i, int j, int k,
                                          const Dims &d, float gamma_lu, int wetting, float cos_t,
                                          float sin_t, float &fx, float &fy, float &fz)
{
    const float eps = 1e-6f;
    if (i < 1 || j < 1 || k < 1 || i > d.nx - 2 || j > d.ny - 2 || k > d.nz - 2) return false;
    const PhiStencil &st = fs.phi;
    float wx = 0.0f, wy = 0.0f, wz = 0.0f;
    bool wall = wetting != 0 && wall_normal(fs.flags, i, j, k, d, wx, wy, wz);
    float kappa, g;
    brackbill_kappa(st, eps, kappa, g);
    if (g < eps) return false;
    float dpx = 0.5f * (st.c[0] - st.c[1]);
    float dpy = 0.5f * (st.c[2] - st.c[3]);
    float dpz = 0.5f * (st.c[4] - st.c[5]);
    fx = gamma_lu * kappa * dpx;
    fy = gamma_lu * kappa * dpy;
    fz = gamma_lu * kappa * dpz;
    if (wall) {
        float cx, cy, cz;
        contact_normal(dpx / g, dpy / g, dpz / g, wx, wy, wz, cos_t, sin_t, cx, cy, cz);
        fx = gamma_lu * kappa * cx * g;
        fy = gamma_lu * kappa * cy * g;
        fz = gamma_lu * kappa * cz * g;
    }
    return true;
}

/* compute_marangoni_force body — reference kernels.py:2187-2217 (cell neither SOLID nor GAS). */
__device__ __forceinline__ bool marangoni_force(const ForceStencil &fs, float fl, float dgamma,
                                                float &fx, float &fy, float &fz)
{
    const float(&pv)[6] = fs.phi.c;
    const float(&tv)[6] = fs.T;
    if (fl < 0.05f) return false;
    float dpx = 0.5f * (pv[0] - pv[1]);
    float dpy = 0.5f * (pv[2] - pv[3]);
    float dpz = 0.5f * (pv[4] - pv[5]);
    float gm = sqrtf(dpx * dpx + dpy * dpy + dpz * dpz);
    if (gm < 1e-6f) return false;
    float nxn = dpx / gm, nyn = dpy / gm, nzn = dpz / gm;
    float dTx = 0.5f * (tv[0] - tv[1]);
    float dTy = 0.5f * (tv[2] - tv[3]);
    float dTz = 0.5f * (tv[4] - tv[5]);
    float ndot = nxn * dTx + nyn * dTy + nzn * dTz;
    float scale = dgamma * gm;
    fx = scale * (dTx - ndot * nxn);
    fy = scale * (dTy - ndot * nyn);
    fz = scale * (dTz - ndot * nzn);
    return true;
}

/* apply_arc_pressure body — reference kernels.py:3070-3083 (cell neither SOLID nor GAS). */
struct ArcPressure {
    float arc_i, arc_j, arc_k, inv2s2, F_peak_lu;
};

/* phi_zp, phi_zm: phi at the clamped k+1 and k-1 cells */
__device__ __forceinline__ bool arc_pressure_force(float phi_zp, float phi_zm, int i, int j, int k,
                                                   const ArcPressure &ap, float &fz)
{
    float dphz = phi_zp - phi_zm;
    if (!(fabsf(dphz) > 0.05f)) return false;
    float di = (float)i - ap.arc_i, dj = (float)j - ap.arc_j, dk = (float)k - ap.arc_k;
    float r2 = di * di + dj * dj + dk * dk * 0.1f;
    fz = -ap.F_peak_lu * waam_expf(-r2 * ap.inv2s2);
    return true;
}

/* vapour recoil on the free surface (reference kernels.py:995-1030, apply_vapor_recoil, and
 * :1052-1090, apply_vapor_recoil_clausius_clapeyron); for a cell that is not SOLID/GAS */
struct Recoil {
    float arc_i, arc_j, inv2s2;
    int clausius;          // 1: p = C_acc P_sat(T) above T_boil; 0: p_peak scaled by T/T_ref
    float F_peak_lu;       // constant model: (p / (rho dx)) dt dt / dx
    float T_ref;           // constant model
    float P_ref, T_boil, L_vap, R_spec, C_acc, dt, dx, rho_ref;  // Clausius-Clapeyron model
};

__device__ __forceinline__ bool recoil_force(float phi_zp, float phi_zm, float T, int i, int j,
                                             const Recoil &r, float &fz)
{
    const float eps = 1e-6f;
    float dphi_z = phi_zp - phi_zm;
    if (fabsf(dphi_z) < 0.05f) return false;
    float di = (float)i - r.arc_i, dj = (float)j - r.arc_j;
    float r2 = di * di + dj * dj;
    if (r.clausius) {
        if (T <= r.T_boil) return false;
        float expo = (r.L_vap / (r.R_spec + eps)) * (1.0f / (r.T_boil + eps) - 1.0f / (T + eps));
        expo = fmn(expo, 12.0f);
        float P_vap = r.C_acc * r.P_ref * waam_expf(expo);
        float F_peak = P_vap / (r.rho_ref * r.dx) * r.dt * r.dt / r.dx;
        fz = -F_peak * waam_expf(-r2 * r.inv2s2);
        return true;
    }
    float t_ratio = fmn(T / (r.T_ref + eps), 3.0f);
    fz = -r.F_peak_lu * t_ratio * waam_expf(-r2 * r.inv2s2);
    return true;
}

/* shielding-gas shear along the surface, away from the arc axis (reference kernels.py:1093-1139,
 * apply_gas_shear_stress); for a cell that is not SOLID/GAS */
struct GasShear {
    float arc_i, arc_j, inv2s2, tau_peak, F_scale;  // F_scale = (1 / (rho dx)) dt dt / dx
};

__device__ __forceinline__ bool gas_shear_force(const PhiStencil &ps, int i, int j,
                                                const GasShear &g, float &fx, float &fy)
{
    const float eps = 1e-6f;
    float dpx = 0.5f * (ps.c[0] - ps.c[1]);
    float dpy = 0.5f * (ps.c[2] - ps.c[3]);
    float dpz = 0.5f * (ps.c[4] - ps.c[5]);
    float gmag = sqrtf(dpx * dpx + dpy * dpy + dpz * dpz);
    if (gmag < 0.08f) return false;
    float di = (float)i - g.arc_i, dj = (float)j - g.arc_j;
    float r2 = di * di + dj * dj;
    float r = sqrtf(r2 + eps);
    float rx = di / r, ry = dj / r;
    float nx_n = dpx / gmag, ny_n = dpy / gmag;
    float rdotn = rx * nx_n + ry * ny_n;
    float tx = rx - rdotn * nx_n, ty = ry - rdotn * ny_n;
    float tmag = sqrtf(tx * tx + ty * ty + eps);
    tx /= tmag;
    ty /= tmag;
    float tau = g.tau_peak * waam_expf(-r2 * g.inv2s2);
    float Fmag = tau * g.F_scale;
    fx = Fmag * tx;
    fy = Fmag * ty;
    return true;
}

}  // namespace waam
