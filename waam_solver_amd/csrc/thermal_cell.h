/*
 * thermal_cell.h — per-cell thermal operators shared by the standalone thermal kernels and the
 * fused thermal kernel.
 */
#pragma once

#include "waam_common.h"

namespace waam {

/* arc_deposition_weight — reference kernels.py:406-437 */
__device__ __forceinline__ float arc_deposition_weight(int k, float arc_k, float fl, float pen,
                                                       int enable)
{
    float w = 1.0f;
    if (enable == 1) {
        if (fl > 0.55f) {
            w = 1.0f;
        } else {
            float dz = arc_k - (float)k;
            if (dz < -0.5f) {
                w = 0.0f;
            } else {
                float depth = fmx(0.0f, dz);
                float inv_pen = 1.0f / (pen + 1e-6f);
                w = waam_expf(-depth * inv_pen);
            }
        }
    }
    return w;
}

/* update_phase — reference kernels.py:1542-1556 (H_sol, H_liq formed in fp32 as there). */
struct PhaseConsts {
    float cp_rho, L_rho, T_s, T_l, H_sol, H_liq;
};

inline PhaseConsts make_phase_consts(float cp_rho, float L_rho, float T_s, float T_l)
{
    PhaseConsts p;
    p.cp_rho = cp_rho;
    p.L_rho = L_rho;
    p.T_s = T_s;
    p.T_l = T_l;
    p.H_sol = cp_rho * T_s;
    p.H_liq = p.H_sol + L_rho;
    return p;
}

__device__ __forceinline__ void phase_of(float h, const PhaseConsts &p, float &T, float &fl)
{
    if (h <= p.H_sol) {
        fl = 0.0f;
        T = h / p.cp_rho;
    } else if (h >= p.H_liq) {
        fl = 1.0f;
        T = p.T_l + (h - p.H_liq) / p.cp_rho;
    } else {
        fl = (h - p.H_sol) / p.L_rho;
        T = p.T_s + fl * (p.T_l - p.T_s);
    }
}

/* clamp_enthalpy_ceiling_scalar — reference kernels.py:1975-1984 */
__device__ __forceinline__ float ceiling_clamp(float h, float H_liq, float cap_l, float cap_s)
{
    if (h > H_liq) {
        if (h > cap_l) h = cap_l;
    } else {
        if (h > cap_s) h = cap_s;
    }
    return h;
}

/* apply_thermal_boundary_losses — reference kernels.py:2055-2082 */
struct LossConsts {
    float T_amb, h_conv, eps_rad, sigma_sb, dt, dx;
    int enable_conv, enable_rad;
};

__device__ __forceinline__ int exposed_faces(const uint8_t *__restrict__ flags, int i, int j,
                                             int k, const Dims &d)
{
    int n = 0;
    if (k + 1 < d.nz && flags[d.idx(i, j, k + 1)] == GAS) n += 1;
    if (j + 1 < d.ny && flags[d.idx(i, j + 1, k)] == GAS) n += 1;
    if (j - 1 >= 0 && flags[d.idx(i, j - 1, k)] == GAS) n += 1;
    if (i + 1 < d.nx && flags[d.idx(i + 1, j, k)] == GAS) n += 1;
    if (i - 1 >= 0 && flags[d.idx(i - 1, j, k)] == GAS) n += 1;
    return n;
}

__device__ __forceinline__ float loss_flux(float T, const LossConsts &lc)
{
    float T_eff = fmn(T, 2500.0f);
    float q = 0.0f;
    if (lc.enable_conv == 1) q += lc.h_conv * (T_eff - lc.T_amb);
    if (lc.enable_rad == 1) {
        float t2 = T_eff * T_eff, a2 = lc.T_amb * lc.T_amb;
        q += lc.eps_rad * lc.sigma_sb * (t2 * t2 - a2 * a2);
    }
    return q;
}

/* _minmod, _masked_at, _limited_upwind_grad — reference kernels.py:1565-1608 */
__device__ __forceinline__ float minmod(float a, float b)
{
    float out = 0.0f;
    if (a * b > 0.0f) out = (fabsf(a) < fabsf(b)) ? a : b;
    return out;
}

__device__ __forceinline__ float masked_at(const float *__restrict__ fld,
                                           const uint8_t *__restrict__ flags, int i, int j, int k,
                                           float fallback, const Dims &d)
{
    int ii = imax(0, imin(i, d.nx - 1));
    int jj = imax(0, imin(j, d.ny - 1));
    int kk = imax(0, imin(k, d.nz - 1));
    size_t c = d.idx(ii, jj, kk);
    float v = fld[c];
    if (flags[c] == GAS) v = fallback;
    return v;
}

__device__ __forceinline__ float upwind_grad(float u, float qmm, float qm, float qc, float qp,
                                             float qpp)
{
    if (u > 0.0f) {
        float d1 = qc - qm;
        float d2 = 0.5f * (3.0f * qc - 4.0f * qm + qmm);
        return minmod(d1, d2);
    }
    float d1 = qp - qc;
    float d2 = 0.5f * (-3.0f * qc + 4.0f * qp - qpp);
    return minmod(d1, d2);
}

/* dH of advect_diffuse_temperature — reference kernels.py:1645-1680, neighbour H from Hs. */
__device__ __forceinline__ float thermal_dH(const float *__restrict__ Hs,
                                            const float *__restrict__ T,
                                            const float *__restrict__ ux,
                                            const float *__restrict__ uy,
                                            const float *__restrict__ uz,
                                            const uint8_t *__restrict__ flags, int i, int j, int k,
                                            uint8_t flag, float alpha_lu, float dt, float cp_rho,
                                            const Dims &d)
{
    size_t c = d.idx(i, j, k);
    float T_c = T[c];
    float T_im = masked_at(T, flags, i - 1, j, k, T_c, d);
    float T_ip = masked_at(T, flags, i + 1, j, k, T_c, d);
    float T_jm = masked_at(T, flags, i, j - 1, k, T_c, d);
    float T_jp = masked_at(T, flags, i, j + 1, k, T_c, d);
    float T_km = masked_at(T, flags, i, j, k - 1, T_c, d);
    float T_kp = masked_at(T, flags, i, j, k + 1, T_c, d);
    float lap = T_ip + T_im + T_jp + T_jm + T_kp + T_km - 6.0f * T_c;
    float dH = cp_rho * alpha_lu * lap * dt;
    if (flag != SOLID) {
        float H_c = Hs[c];
        float u = ux[c], v = uy[c], w = uz[c];
        float H_im = masked_at(Hs, flags, i - 1, j, k, H_c, d);
        float H_ip = masked_at(Hs, flags, i + 1, j, k, H_c, d);
        float H_imm = masked_at(Hs, flags, i - 2, j, k, H_im, d);
        float H_ipp = masked_at(Hs, flags, i + 2, j, k, H_ip, d);
        float H_jm = masked_at(Hs, flags, i, j - 1, k, H_c, d);
        float H_jp = masked_at(Hs, flags, i, j + 1, k, H_c, d);
        float H_jmm = masked_at(Hs, flags, i, j - 2, k, H_jm, d);
        float H_jpp = masked_at(Hs, flags, i, j + 2, k, H_jp, d);
        float H_km = masked_at(Hs, flags, i, j, k - 1, H_c, d);
        float H_kp = masked_at(Hs, flags, i, j, k + 1, H_c, d);
        float H_kmm = masked_at(Hs, flags, i, j, k - 2, H_km, d);
        float H_kpp = masked_at(Hs, flags, i, j, k + 2, H_kp, d);
        float gx = upwind_grad(u, H_imm, H_im, H_c, H_ip, H_ipp);
        float gy = upwind_grad(v, H_jmm, H_jm, H_c, H_jp, H_jpp);
        float gz = upwind_grad(w, H_kmm, H_km, H_c, H_kp, H_kpp);
        dH -= (u * gx + v * gy + w * gz) * dt;
    }
    return dH;
}

}  // namespace waam
