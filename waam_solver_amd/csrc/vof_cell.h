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
    /* every operand is loaded (at clamped indices) before any is selected, so the loads of
     * all faces issue together instead of as a flags -> velocity -> phi chain */
    size_t c0 = d.idx(ii0, imax(0, imin(j0, d.ny - 1)), imax(0, imin(k0, d.nz - 1)));
    size_t c1 = d.idx(ii1, imax(0, imin(j1, d.ny - 1)), imax(0, imin(k1, d.nz - 1)));
    uint8_t f0 = flags[c0], f1 = flags[c1];
    float v0 = vel[c0], v1 = vel[c1], p0 = phi[c0], p1 = phi[c1];
    float u0 = f0 != GAS ? v0 : 0.0f;
    float u1 = f1 != GAS ? v1 : 0.0f;
    float uf = 0.5f * (u0 + u1);
    float fx = uf > 0.0f ? uf * p0 : uf * p1;
    /* selects, not branches: a branch would let the compiler sink the loads behind it */
    flux = (in0 && in1 && f0 != SOLID && f1 != SOLID) ? fx : flux;
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
    float pc = src[c];
    float fe = vof_face_flux(src, ux, flags, i, j, k, i + 1, j, k, d);
    float fw = vof_face_flux(src, ux, flags, i - 1, j, k, i, j, k, d);
    float fn = vof_face_flux(src, uy, flags, i, j, k, i, j + 1, k, d);
    float fs = vof_face_flux(src, uy, flags, i, j - 1, k, i, j, k, d);
    float ft = vof_face_flux(src, uz, flags, i, j, k, i, j, k + 1, d);
    float fb = vof_face_flux(src, uz, flags, i, j, k - 1, i, j, k, d);
    float r = pc - (fe - fw) - (fn - fs) - (ft - fb);
    return flags[c] == SOLID ? pc : r;  // SOLID copies through
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
    /* all operands first (clamped indices), then the rules as selects on them */
    size_t c = d.idx(i, j, k);
    size_t cb = d.idx(i, j, imax(k - 1, 0)), cbb = d.idx(i, j, imax(k - 2, 0));
    float p0 = pin[c], pb0 = pin[cb];
    uint8_t f = flags[c], fb = flags[cb], fbb = flags[cbb];
    float p_jm = pin[d.idx(i, imax(j - 1, 0), k)], p_jp = pin[d.idx(i, imin(j + 1, d.ny - 1), k)];
    float p_im = pin[d.idx(imax(i - 1, 0), j, k)], p_ip = pin[d.idx(imin(i + 1, d.nx - 1), j, k)];
    float p = p0;
    float film = 0.5f * (1.0f - cos_t);
    /* rule 2 of the cell below (a liquid cell on SOLID raises the GAS cell above it) */
    float pb = (fb == GAS && pb0 < film) ? film : pb0;
    float g = fmx(film, fmn(pb - sin_t, 0.98f));
    p = (k >= 2 && f == GAS && fbb == SOLID && fb != SOLID && pb > 0.55f && p < g) ? g : p;
    /* rule 1: GAS directly above SOLID holds at least the film */
    bool below_solid = k > 0 && fb == SOLID;
    p = (below_solid && f == GAS && p < film) ? film : p;
    /* rule 3: GAS above a mostly-liquid cell */
    float t = fmx(film, fmn(pb0 - sin_t, 0.95f));
    p = (k > 0 && pb0 > 0.55f && f == GAS && p < t) ? t : p;
    /* rule 4: lateral spreading into a mostly-empty cell (the 0.45 test sees p before it) */
    bool r4 = p < 0.45f;
    float spread = film * (0.85f + 0.15f * sin_t);
    p = (r4 && j - 1 >= 0 && p_jm > 0.55f && p < spread) ? spread : p;
    p = (r4 && j + 1 < d.ny && p_jp > 0.55f && p < spread) ? spread : p;
    p = (r4 && i - 1 >= 0 && p_im > 0.55f && p < spread) ? spread : p;
    p = (r4 && i + 1 < d.nx && p_ip > 0.55f && p < spread) ? spread : p;
    /* rule 5: pin wetted cells on SOLID */
    p = (below_solid && p > 0.2f && p < 0.65f) ? 0.65f : p;
    return f == SOLID ? p0 : p;
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
