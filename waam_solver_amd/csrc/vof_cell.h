/*
 * vof_cell.h — per-cell VOF operators shared by the standalone and fused VOF kernels.
 */
#pragma once

#include "waam_common.h"

namespace waam {

/* _vof_face_flux — reference kernels.py:2275-2317 (x periodic, y/z no-flux walls, no flux
 * across a SOLID face, GAS velocity taken as zero). */
__device__ __forceinline__ float vof_face_flux(const float *__restrict__ phi,
                                               const float *__restrict__ vel,
                                               const uint8_t *__restrict__ flags, int i0, int j0,
                                               int k0, int i1, int j1, int k1, const Dims &d)
{
    float flux = 0.0f;
    int ii0 = i0 < 0 ? i0 + d.nx : (i0 >= d.nx ? i0 - d.nx : i0);
    int ii1 = i1 < 0 ? i1 + d.nx : (i1 >= d.nx ? i1 - d.nx : i1);
    bool in0 = j0 >= 0 && k0 >= 0 && j0 < d.ny && k0 < d.nz;
    bool in1 = j1 >= 0 && k1 >= 0 && j1 < d.ny && k1 < d.nz;
    if (in0 && in1) {
        size_t c0 = d.idx(ii0, j0, k0), c1 = d.idx(ii1, j1, k1);
        uint8_t f0 = flags[c0], f1 = flags[c1];
        if (f0 != SOLID && f1 != SOLID) {
            float u0 = 0.0f, u1 = 0.0f;
            if (f0 != GAS) u0 = vel[c0];
            if (f1 != GAS) u1 = vel[c1];
            float uf = 0.5f * (u0 + u1);
            flux = uf > 0.0f ? uf * phi[c0] : uf * phi[c1];
        }
    }
    return flux;
}

/* advect_phi body — reference kernels.py:2342-2361 */
__device__ __forceinline__ float advected_phi(const float *__restrict__ src,
                                              const float *__restrict__ ux,
                                              const float *__restrict__ uy,
                                              const float *__restrict__ uz,
                                              const uint8_t *__restrict__ flags, int i, int j,
                                              int k, const Dims &d)
{
    size_t c = d.idx(i, j, k);
    if (flags[c] == SOLID) return src[c];
    float fe = vof_face_flux(src, ux, flags, i, j, k, i + 1, j, k, d);
    float fw = vof_face_flux(src, ux, flags, i - 1, j, k, i, j, k, d);
    float fn = vof_face_flux(src, uy, flags, i, j, k, i, j + 1, k, d);
    float fs = vof_face_flux(src, uy, flags, i, j - 1, k, i, j, k, d);
    float ft = vof_face_flux(src, uz, flags, i, j, k, i, j, k + 1, d);
    float fb = vof_face_flux(src, uz, flags, i, j, k - 1, i, j, k, d);
    return src[c] - (fe - fw) - (fn - fs) - (ft - fb);
}

/* reinitialize_phi — reference kernels.py:2380-2391 */
__device__ __forceinline__ float reinit_phi(float p, uint8_t f)
{
    if (f == SOLID) return 1.0f;
    if (f == GAS) return p < 1e-6f ? 0.0f : fmn(1.0f, p);
    return fmx(0.0f, fmn(1.0f, p));
}

/* apply_contact_angle_phi_bc — reference kernels.py:949-991 for one cell, Jacobi form: all
 * neighbour phi from `pin`; the reference's one cross-cell write (rule 2: a liquid cell sitting
 * on SOLID raises the GAS cell above it) is evaluated here by that upper cell and applied before
 * its own rules, the order a lexicographic sweep applies it in. */
__device__ __forceinline__ float contact_bc_phi(const float *__restrict__ pin,
                                                const uint8_t *__restrict__ flags, int i, int j,
                                                int k, float cos_t, float sin_t, const Dims &d)
{
    size_t c = d.idx(i, j, k);
    float p = pin[c];
    uint8_t f = flags[c];
    if (f == SOLID) return p;
    float film = 0.5f * (1.0f - cos_t);
    if (k >= 2 && f == GAS) {
        uint8_t fb = flags[c - 1], fbb = flags[c - 2];
        if (fbb == SOLID && fb != SOLID) {
            float pb = pin[c - 1];
            if (fb == GAS && pb < film) pb = film;
            if (pb > 0.55f) {
                float g = fmx(film, fmn(pb - sin_t, 0.98f));
                if (p < g) p = g;
            }
        }
    }
    bool below_solid = k > 0 && flags[c - 1] == SOLID;
    if (below_solid && f == GAS) {
        if (p < film) p = film;
    }
    if (k > 0 && pin[c - 1] > 0.55f && f == GAS) {
        float t = fmx(film, fmn(pin[c - 1] - sin_t, 0.95f));
        if (p < t) p = t;
    }
    if (p < 0.45f) {
        float spread = film * (0.85f + 0.15f * sin_t);
        if (j - 1 >= 0 && pin[d.idx(i, j - 1, k)] > 0.55f && p < spread) p = spread;
        if (j + 1 < d.ny && pin[d.idx(i, j + 1, k)] > 0.55f && p < spread) p = spread;
        if (i - 1 >= 0 && pin[d.idx(i - 1, j, k)] > 0.55f && p < spread) p = spread;
        if (i + 1 < d.nx && pin[d.idx(i + 1, j, k)] > 0.55f && p < spread) p = spread;
    }
    if (below_solid && p > 0.2f) {
        if (p < 0.65f) p = 0.65f;
    }
    return p;
}

/* update_flags_from_phi — reference kernels.py:2406-2419 */
__device__ __forceinline__ uint8_t classify_cell(int k, int nz_solid, float p, float fl)
{
    if (k < nz_solid) return SOLID;
    if (p > 0.85f && fl < 0.05f) return SOLID;
    if (p > 0.55f && fl > 0.1f) return FLUID;
    if (p < 0.05f) return GAS;
    return IFACE;
}

}  // namespace waam
