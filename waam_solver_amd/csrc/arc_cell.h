/*
 * arc_cell.h — arc heat-source weights (reference kernels.py:406-622) and the compensated block
 * reduction used for the weight total; shared by the standalone and fused arc kernels.
 */
#pragma once

#include "thermal_cell.h"

namespace waam {

struct ArcParams {
    int kind;  // 0 Gaussian (Gaussian2D / ConicalVolume), 1 Goldak
    float arc_i, arc_j, arc_k;
    float inv2s2;                                  // Gaussian
    float sgn, ff, fr, af, ar, b, cax;             // Goldak
    float pen;
    int enable;
};

/* Returns false for cells the reference skips (w_dep < 1e-6); otherwise the shape weight wg
 * (Gaussian exp or Goldak pdf) and the surface weight wdep, whose product is the pass-1 term and
 * (epw*wg)*wdep the pass-2 increment. */
__device__ __forceinline__ bool arc_cell_weights(const ArcParams &p, int i, int j, int k, float fl,
                                                 float &wg, float &wdep)
{
    wdep = arc_deposition_weight(k, p.arc_k, fl, p.pen, p.enable);
    if (wdep < 1e-6f) return false;
    float di = (float)i - p.arc_i, dj = (float)j - p.arc_j, dk = (float)k - p.arc_k;
    if (p.kind == 0) {
        float r2 = di * di + dj * dj + dk * dk * 0.25f;
        wg = waam_expf(-r2 * p.inv2s2);
        return true;
    }
    const float eps = 1e-6f;
    float x = di * p.sgn;
    bool front = x >= 0.0f;
    float a = front ? p.af : p.ar;
    float frac = front ? p.ff : p.fr;
    wg = frac * waam_expf(-3.0f * ((x * x) / (a * a + eps) + (dj * dj) / (p.b * p.b + eps) +
                                   (dk * dk) / (p.cax * p.cax + eps)));
    return true;
}

__device__ __forceinline__ void block_reduce_pair(double &s, double &c, double *sh_s, double *sh_c)
{
    int t = threadIdx.x;
    sh_s[t] = s;
    sh_c[t] = c;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (t < off) {
            double a = sh_s[t], b = sh_s[t + off], ss, e;
            two_sum(a, b, ss, e);
            sh_s[t] = ss;
            sh_c[t] = sh_c[t] + sh_c[t + off] + e;
        }
        __syncthreads();
    }
    s = sh_s[0];
    c = sh_c[0];
}


}  // namespace waam
