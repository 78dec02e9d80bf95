/*
 * lbm_cell.h — per-cell D3Q19 collision operators (device functions shared by the standalone
 * collide kernels and the fused collide-stream kernel).
 *
 * mrt_cell: two-rate central-moment collision, reference cumulant_kernel.py:59-209.
 * srt_cell: BGK with Guo forcing and Carman-Kozeny drag, reference kernels.py:2815-2885.
 * Both evaluate the reference's expressions in its operand order (fp32, no contraction);
 * products of a float with a lattice direction component (-1, 0, 1) are kept as products so
 * that the result matches the reference term for term, including for non-finite inputs.
 */
#pragma once

#include "waam_common.h"

namespace waam {

struct CellOut {
    float f[Q];
    float rho, u, v, w;
};

/* Weights of the hand-expanded reference reconstruction (truncated decimals,
 * cumulant_kernel.py:147-202), rounded from double like every Python literal. */
__device__ __forceinline__ constexpr float w_mrt(int q)
{
    return q == 0 ? (float)0.3333333333 : (q < 7 ? (float)0.0555555556 : (float)0.0277777778);
}

__device__ __forceinline__ void mrt_cell(const float (&f)[Q], float fx, float fy, float fz,
                                         float fl, float omega_s, float omega_b, float C_darcy,
                                         CellOut &o)
{
    const float cs2 = (float)(1.0 / 3.0);
    const float inv_cs2 = 3.0f;
    const float eps = 1.0e-6f;
    float r = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] +
              f[11] + f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
    float inv_rho = 1.0f / (r + eps);
    float jx = f[1] * 1.0f + f[2] * -1.0f + f[7] * 1.0f + f[8] * -1.0f + f[9] * 1.0f +
               f[10] * -1.0f + f[11] * 1.0f + f[12] * -1.0f + f[13] * 1.0f + f[14] * -1.0f;
    float jy = f[3] * 1.0f + f[4] * -1.0f + f[7] * 1.0f + f[8] * -1.0f + f[9] * -1.0f +
               f[10] * 1.0f + f[15] * 1.0f + f[16] * -1.0f + f[17] * 1.0f + f[18] * -1.0f;
    float jz = f[5] * 1.0f + f[6] * -1.0f + f[11] * 1.0f + f[12] * -1.0f + f[13] * -1.0f +
               f[14] * 1.0f + f[15] * 1.0f + f[16] * -1.0f + f[17] * -1.0f + f[18] * 1.0f;
    float uxr = (jx + 0.5f * fx) * inv_rho;
    float uyr = (jy + 0.5f * fy) * inv_rho;
    float uzr = (jz + 0.5f * fz) * inv_rho;
    float one_m = 1.0f - fl;
    float ck = 1.0f + C_darcy * (one_m * one_m) / (fl * (fl * fl) + eps);
    float u = uxr / ck, v = uyr / ck, w = uzr / ck;

    float Pxx = f[1] + f[2] + f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] + f[14];
    float Pyy = f[3] + f[4] + f[7] + f[8] + f[9] + f[10] + f[15] + f[16] + f[17] + f[18];
    float Pzz = f[5] + f[6] + f[11] + f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
    float Pxy = f[7] + f[8] * -1.0f * -1.0f + f[9] * -1.0f + f[10] * -1.0f;
    float Pxz = f[11] + f[12] * -1.0f * -1.0f + f[13] * -1.0f + f[14] * -1.0f;
    float Pyz = f[15] + f[16] * -1.0f * -1.0f + f[17] * -1.0f + f[18] * -1.0f;

    float kxx = Pxx - r * u * u;
    float kyy = Pyy - r * v * v;
    float kzz = Pzz - r * w * w;
    float kxy = Pxy - r * u * v;
    float kxz = Pxz - r * u * w;
    float kyz = Pyz - r * v * w;

    float kmean = (kxx + kyy + kzz) / 3.0f;
    float kmean_p = kmean - omega_b * (kmean - r * cs2);
    float kxxp = kmean_p + (1.0f - omega_s) * (kxx - kmean);
    float kyyp = kmean_p + (1.0f - omega_s) * (kyy - kmean);
    float kzzp = kmean_p + (1.0f - omega_s) * (kzz - kmean);
    float kxyp = kxy - omega_s * kxy;
    float kxzp = kxz - omega_s * kxz;
    float kyzp = kyz - omega_s * kyz;

    float u2 = u * u + v * v + w * w;
    float uF = u * fx + v * fy + w * fz;
    float rcs = r * cs2;
    float dxx = kxxp - rcs, dyy = kyyp - rcs, dzz = kzzp - rcs;
    float s_pref = 1.0f - 0.5f * omega_s;
    /* Terms shared by all directions, in the reference's association:
     *   r*3*eu, ((r*3)*3)*0.5*eu*eu, ((r*0.5)*3)*u2, 4.5*K, (w*(1-0.5 om))*(3 eF + 9 eF eu - 3 uF) */
    float r3 = r * inv_cs2;
    float r9h = r * inv_cs2 * inv_cs2 * 0.5f;
    float ru2 = r * 0.5f * inv_cs2 * u2;
    const float k45 = 0.5f * inv_cs2 * inv_cs2;
    float uF3 = inv_cs2 * (uF);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const float exq = (float)ex(q), eyq = (float)ey(q), ezq = (float)ez(q);
        float eu = exq * u + eyq * v + ezq * w;
        float eF = exq * fx + eyq * fy + ezq * fz;
        float K;
        switch (q) {
            case 0: K = 0.0f; break;
            case 1: case 2: K = dxx; break;
            case 3: case 4: K = dyy; break;
            case 5: case 6: K = dzz; break;
            case 7: case 8: K = (dxx + dyy) + (2.0f * kxyp); break;
            case 9: case 10: K = (dxx + dyy) + (-2.0f * kxyp); break;
            case 11: case 12: K = (dxx + dzz) + (2.0f * kxzp); break;
            case 13: case 14: K = (dxx + dzz) + (-2.0f * kxzp); break;
            case 15: case 16: K = (dyy + dzz) + (2.0f * kyzp); break;
            default: K = (dyy + dzz) + (-2.0f * kyzp); break;
        }
        const float wqq = w_mrt(q);
        float feq = wqq * (r + r3 * (eu) + r9h * (eu) * (eu) - ru2 + k45 * (K));
        float S = wqq * s_pref * (inv_cs2 * (eF) + inv_cs2 * inv_cs2 * (eF) * (eu) - uF3);
        o.f[q] = feq + S;
    }
    o.rho = r;
    o.u = u;
    o.v = v;
    o.w = w;
}

__device__ __forceinline__ void srt_cell(const float (&f)[Q], float fx, float fy, float fz,
                                         float fl, float omega, float C_darcy, CellOut &o)
{
    const float eps = 1.0e-6f;
    float r = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        r += f[q];
        jx += f[q] * (float)ex(q);
        jy += f[q] * (float)ey(q);
        jz += f[q] * (float)ez(q);
    }
    float inv_r = 1.0f / (r + eps);
    float uxr = jx * inv_r + 0.5f * fx * inv_r;
    float uyr = jy * inv_r + 0.5f * fy * inv_r;
    float uzr = jz * inv_r + 0.5f * fz * inv_r;
    float one_m = 1.0f - fl;
    float ck = 1.0f + C_darcy * (one_m * one_m) / (fl * (fl * fl) + eps);
    float u = uxr / ck, v = uyr / ck, w = uzr / ck;
    float u2 = u * u + v * v + w * w;
    float uF3 = 3.0f * (u * fx + v * fy + w * fz);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const float exq = (float)ex(q), eyq = (float)ey(q), ezq = (float)ez(q);
        float eu = exq * u + eyq * v + ezq * w;
        float feq = wq(q) * r * (1.0f + 3.0f * eu + 4.5f * eu * eu - 1.5f * u2);
        float Fe = exq * fx + eyq * fy + ezq * fz;
        float S = wq(q) * (1.0f - 0.5f * omega) * (3.0f * (Fe) + 9.0f * (eu * Fe) - uF3);
        o.f[q] = f[q] - omega * (f[q] - feq) + S;
    }
    o.rho = r;
    o.u = u;
    o.v = v;
    o.w = w;
}

}  // namespace waam
