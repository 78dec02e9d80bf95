/*
 * forces_cell.h — per-cell body-force terms shared by the standalone force kernels and the
 * fused collide kernel. Each returns the increment the reference adds with += so a caller can
 * accumulate terms in the reference's order (coupled_step.py:298-362).
 */
#pragma once

#include "waam_common.h"

namespace waam {

/* _phi_unit_normal_at — reference kernels.py:769-796 */
__device__ __forceinline__ void phi_normal(const float *__restrict__ phi, int i, int j, int k,
                                           const Dims &d, float eps, float &nx_, float &ny_,
                                           float &nz_, float &g)
{
    float dpx = 0.5f * (phi[d.idx(imin(i + 1, d.nx - 1), j, k)] - phi[d.idx(imax(i - 1, 0), j, k)]);
    float dpy = 0.5f * (phi[d.idx(i, imin(j + 1, d.ny - 1), k)] - phi[d.idx(i, imax(j - 1, 0), k)]);
    float dpz = 0.5f * (phi[d.idx(i, j, imin(k + 1, d.nz - 1))] - phi[d.idx(i, j, imax(k - 1, 0))]);
    g = sqrtf(dpx * dpx + dpy * dpy + dpz * dpz);
    nx_ = 0.0f;
    ny_ = 0.0f;
    nz_ = 0.0f;
    if (g >= eps) {
        nx_ = dpx / g;
        ny_ = dpy / g;
        nz_ = dpz / g;
    }
}

/* _brackbill_curvature_at — reference kernels.py:799-855 */
__device__ __forceinline__ void brackbill_kappa(const float *__restrict__ phi, int i, int j, int k,
                                                const Dims &d, float eps, float &kappa, float &g0)
{
    float n0x, n0y, n0z;
    phi_normal(phi, i, j, k, d, eps, n0x, n0y, n0z, g0);
    kappa = 0.0f;
    if (g0 >= eps) {
        float a, b, nxp, nxm, nyp, nym, nzp, nzm, gp, gm, gp2, gm2, gp3, gm3;
        phi_normal(phi, imin(i + 1, d.nx - 1), j, k, d, eps, nxp, a, b, gp);
        phi_normal(phi, imax(i - 1, 0), j, k, d, eps, nxm, a, b, gm);
        phi_normal(phi, i, imin(j + 1, d.ny - 1), k, d, eps, a, nyp, b, gp2);
        phi_normal(phi, i, imax(j - 1, 0), k, d, eps, a, nym, b, gm2);
        phi_normal(phi, i, j, imin(k + 1, d.nz - 1), d, eps, a, b, nzp, gp3);
        phi_normal(phi, i, j, imax(k - 1, 0), d, eps, a, b, nzm, gm3);
        if (gp < eps) nxp = n0x;
        if (gm < eps) nxm = n0x;
        if (gp2 < eps) nyp = n0y;
        if (gm2 < eps) nym = n0y;
        if (gp3 < eps) nzp = n0z;
        if (gm3 < eps) nzm = n0z;
        float div_n = 0.5f * (nxp - nxm) + 0.5f * (nyp - nym) + 0.5f * (nzp - nzm);
        kappa = -div_n;
    }
}

/* _solid_wall_normal — reference kernels.py:680-719 */
__device__ __forceinline__ bool wall_normal(const uint8_t *__restrict__ flags, int i, int j, int k,
                                            const Dims &d, float &wx, float &wy, float &wz)
{
    wx = 0.0f;
    wy = 0.0f;
    wz = 0.0f;
    bool has = false;
    if (i > 0 && flags[d.idx(i - 1, j, k)] == SOLID) { wx += 1.0f; has = true; }
    if (i < d.nx - 1 && flags[d.idx(i + 1, j, k)] == SOLID) { wx -= 1.0f; has = true; }
    if (j > 0 && flags[d.idx(i, j - 1, k)] == SOLID) { wy += 1.0f; has = true; }
    if (j < d.ny - 1 && flags[d.idx(i, j + 1, k)] == SOLID) { wy -= 1.0f; has = true; }
    if (k > 0 && flags[d.idx(i, j, k - 1)] == SOLID) { wz += 1.0f; has = true; }
    if (k < d.nz - 1 && flags[d.idx(i, j, k + 1)] == SOLID) { wz -= 1.0f; has = true; }
    float m = sqrtf(wx * wx + wy * wy + wz * wz);
    if (m > 1e-8f) {
        wx /= m;
        wy /= m;
        wz /= m;
    }
    return has;
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

/* compute_csf_tension body — reference kernels.py:885-926. Returns false when the reference
 * adds nothing (rim, |grad phi| < eps). Caller has checked the cell is neither SOLID nor GAS. */
__device__ __forceinline__ bool csf_force(const float *__restrict__ phi,
                                          const uint8_t *__restrict__ flags, int i, int j, int k,
                                          const Dims &d, float gamma_lu, int wetting, float cos_t,
                                          float sin_t, float &fx, float &fy, float &fz)
{
    const float eps = 1e-6f;
    if (i < 1 || j < 1 || k < 1 || i > d.nx - 2 || j > d.ny - 2 || k > d.nz - 2) return false;
    float kappa, g;
    brackbill_kappa(phi, i, j, k, d, eps, kappa, g);
    if (g < eps) return false;
    float dpx = 0.5f * (phi[d.idx(imin(i + 1, d.nx - 1), j, k)] - phi[d.idx(imax(i - 1, 0), j, k)]);
    float dpy = 0.5f * (phi[d.idx(i, imin(j + 1, d.ny - 1), k)] - phi[d.idx(i, imax(j - 1, 0), k)]);
    float dpz = 0.5f * (phi[d.idx(i, j, imin(k + 1, d.nz - 1))] - phi[d.idx(i, j, imax(k - 1, 0))]);
    fx = gamma_lu * kappa * dpx;
    fy = gamma_lu * kappa * dpy;
    fz = gamma_lu * kappa * dpz;
    if (wetting != 0) {
        float wx, wy, wz;
        if (wall_normal(flags, i, j, k, d, wx, wy, wz)) {
            float cx, cy, cz;
            contact_normal(dpx / g, dpy / g, dpz / g, wx, wy, wz, cos_t, sin_t, cx, cy, cz);
            fx = gamma_lu * kappa * cx * g;
            fy = gamma_lu * kappa * cy * g;
            fz = gamma_lu * kappa * cz * g;
        }
    }
    return true;
}

/* compute_marangoni_force body — reference kernels.py:2187-2217 (cell neither SOLID nor GAS). */
__device__ __forceinline__ bool marangoni_force(const float *__restrict__ T,
                                                const float *__restrict__ phi, float fl, int i,
                                                int j, int k, const Dims &d, float dgamma,
                                                float &fx, float &fy, float &fz)
{
    if (fl < 0.05f) return false;
    int ip = imin(i + 1, d.nx - 1), im = imax(i - 1, 0);
    int jp = imin(j + 1, d.ny - 1), jm = imax(j - 1, 0);
    int kp = imin(k + 1, d.nz - 1), km = imax(k - 1, 0);
    float dpx = 0.5f * (phi[d.idx(ip, j, k)] - phi[d.idx(im, j, k)]);
    float dpy = 0.5f * (phi[d.idx(i, jp, k)] - phi[d.idx(i, jm, k)]);
    float dpz = 0.5f * (phi[d.idx(i, j, kp)] - phi[d.idx(i, j, km)]);
    float gm = sqrtf(dpx * dpx + dpy * dpy + dpz * dpz);
    if (gm < 1e-6f) return false;
    float nxn = dpx / gm, nyn = dpy / gm, nzn = dpz / gm;
    float dTx = 0.5f * (T[d.idx(ip, j, k)] - T[d.idx(im, j, k)]);
    float dTy = 0.5f * (T[d.idx(i, jp, k)] - T[d.idx(i, jm, k)]);
    float dTz = 0.5f * (T[d.idx(i, j, kp)] - T[d.idx(i, j, km)]);
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

__device__ __forceinline__ bool arc_pressure_force(const float *__restrict__ phi, int i, int j,
                                                   int k, const Dims &d, const ArcPressure &ap,
                                                   float &fz)
{
    float dphz = phi[d.idx(i, j, imin(k + 1, d.nz - 1))] - phi[d.idx(i, j, imax(k - 1, 0))];
    if (!(fabsf(dphz) > 0.05f)) return false;
    float di = (float)i - ap.arc_i, dj = (float)j - ap.arc_j, dk = (float)k - ap.arc_k;
    float r2 = di * di + dj * dj + dk * dk * 0.1f;
    fz = -ap.F_peak_lu * waam_expf(-r2 * ap.inv2s2);
    return true;
}

}  // namespace waam
