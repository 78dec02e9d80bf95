/*
 * waam_oracle.c — CPU restatement of the waam_twin coupled-step kernels.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP implementation in
 * waam_solver_amd/csrc; it is compiled into oracle/_build/liboracle*.so and loaded only by
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg. Product code never links it.
 *
 * Every function restates one Taichi kernel of the reference (file:line given per function)
 * in the reference's own expression order, in fp32, with FMA contraction disabled
 * (-ffp-contract=off) so each + - * / sqrt is one correctly rounded IEEE operation.
 * Taichi semantics pinned here (Taichi itself cannot run in this environment):
 *   - struct-for over a dense (nx,ny,nz) field visits i, then j, then k fastest;
 *     "serial" kernels below use that order and update in place;
 *   - Python literals are fp32 constants; int*float promotes the int to fp32;
 *   - x**2 = x*x, x**3 = x*(x*x), x**4 = (x*x)*(x*x) (exponentiation by squaring);
 *   - ti.math.exp is waam_expf (waam_math.h), shared with the GPU code so that both
 *     evaluate it identically; cos/sin of the contact angle are passed in precomputed;
 *   - float atomics that accumulate a sum (arc weight total) are a compensated double sum
 *     rounded once to fp32; integer-valued sums are exact.
 * Layout: every scalar field is C-order (nx,ny,nz) with k fastest; populations are
 * (19,nx,ny,nz); flags are uint8 with FLUID=0, SOLID=1, GAS=2, IFACE=4 (grid.py:63-66).
 *
 * Kernels whose reference form reads neighbours it also writes (in-place, order dependent)
 * come in two variants: *_serial (lexicographic in-place order, the ti.cpu single-thread
 * behaviour) and *_jacobi (all neighbour reads from the pre-kernel state), the latter being
 * the semantics the GPU implements.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../waam_solver_amd/csrc/waam_math.h"

#ifdef ORACLE_F64
/* The same kernels in binary64: every fp32 variable, argument and cast below becomes a double
 * and the fp32 math calls their double forms (exp for the shared fp32 exponential). Literals
 * written as fp32 constants keep their fp32 values; the operation sequence is unchanged. */
#define float double
#define sqrtf sqrt
#define fabsf fabs
#define fmaxf fmax
#define fminf fmin
#define waam_expf exp
#endif

#define FLUID 0
#define SOLID 1
#define GAS 2
#define IFACE 4
/* the moving lid of the cavity case (not a reference flag) */
#define CAVITY_LID 3
#define Q 19

/* D3Q19 velocity set, weights and opposites (grid.py:19-34, kernels.py:42-51). */
static const int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
static const int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
static const float WQ[Q] = {
    (float)(1.0 / 3.0),
    (float)(1.0 / 18.0), (float)(1.0 / 18.0), (float)(1.0 / 18.0),
    (float)(1.0 / 18.0), (float)(1.0 / 18.0), (float)(1.0 / 18.0),
    (float)(1.0 / 36.0), (float)(1.0 / 36.0), (float)(1.0 / 36.0), (float)(1.0 / 36.0),
    (float)(1.0 / 36.0), (float)(1.0 / 36.0), (float)(1.0 / 36.0), (float)(1.0 / 36.0),
    (float)(1.0 / 36.0), (float)(1.0 / 36.0), (float)(1.0 / 36.0), (float)(1.0 / 36.0)};

#define IDX(i, j, k) ((((size_t)(i)) * (size_t)ny + (size_t)(j)) * (size_t)nz + (size_t)(k))
#define NCELL ((size_t)nx * (size_t)ny * (size_t)nz)

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline float fminf_(float a, float b) { return a < b ? a : b; }
static inline float fmaxf_(float a, float b) { return a > b ? a : b; }
/* ti.min/ti.max on floats: a < b ? a : b (no NaN special-casing needed on these paths). */

int oracle_abi_version(void) { return 1; }

/* ------------------------------------------------------------------------------------------
 * Initialisation — init_grid kernels.py:62-120
 * ---------------------------------------------------------------------------------------- */
void or_init_grid(float *f, float *rho, float *ux, float *uy, float *uz, float *T, float *H,
                  float *f_l, float *phi, uint8_t *flags, float T_ambient, float rho0,
                  float cp_rho, int nz_solid, int pi0, int pi1, int pj0, int pj1, int nx,
                  int ny, int nz)
{
    size_t N = NCELL;
    for (int q = 0; q < Q; ++q) {
        float v = WQ[q] * rho0;
        for (size_t c = 0; c < N; ++c) f[(size_t)q * N + c] = v;
    }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                rho[c] = rho0;
                ux[c] = 0.0f;
                uy[c] = 0.0f;
                uz[c] = 0.0f;
                T[c] = T_ambient;
                H[c] = cp_rho * T_ambient;
                f_l[c] = 0.0f;
                if (nz_solid < 0) {
                    flags[c] = FLUID;
                    phi[c] = 1.0f;
                    f_l[c] = 1.0f;
                } else {
                    int in_plate = (i >= pi0 && i < pi1 && j >= pj0 && j < pj1);
                    if (in_plate && k < nz_solid) {
                        phi[c] = 1.0f;
                        flags[c] = SOLID;
                    } else {
                        phi[c] = 0.0f;
                        flags[c] = GAS;
                    }
                }
            }
}

/* ------------------------------------------------------------------------------------------
 * Arc heat — arc_deposition_weight kernels.py:406-437, inject_arc_heat :443-510,
 * _goldak_pdf_weight :513-545, inject_goldak_heat :548-622.
 * The reference accumulates the weight total with a float atomic; here it is a double sum
 * rounded once to fp32 (order independent up to double rounding).
 * ---------------------------------------------------------------------------------------- */
/* Compensated (TwoSum) accumulation: the returned double is the exact sum of the terms to
 * within a couple of double ulps, whatever the order, so rounding it to fp32 is order
 * independent in practice. */
typedef struct { double s, c; } csum_t;
static inline void csum_add(csum_t *a, double x)
{
    double t = a->s + x;
    double bb = t - a->s;
    a->c += (a->s - (t - bb)) + (x - bb);
    a->s = t;
}

static inline float arc_deposition_weight(int k, float arc_k, float fl, float pen, int enable)
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
                float depth = fmaxf_(0.0f, dz);
                float inv_pen = 1.0f / (pen + 1e-6f);
                w = waam_expf(-depth * inv_pen);
            }
        }
    }
    return w;
}

float or_arc_weight_sum(const uint8_t *flags, const float *f_l, float arc_i, float arc_j,
                        float arc_k, float sigma, float pen, int enable, int nx, int ny, int nz)
{
    float inv2s2 = 1.0f / (2.0f * sigma * sigma);
    csum_t acc = {0.0, 0.0};
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == GAS) continue;
                float w_dep = arc_deposition_weight(k, arc_k, f_l[c], pen, enable);
                if (w_dep < 1e-6f) continue;
                float di = (float)i - arc_i, dj = (float)j - arc_j, dk = (float)k - arc_k;
                float r2 = di * di + dj * dj + dk * dk * 0.25f;
                csum_add(&acc, (double)(waam_expf(-r2 * inv2s2) * w_dep));
            }
    return (float)(acc.s + acc.c);
}

void or_inject_arc_heat(float *H, const uint8_t *flags, const float *f_l, float *norm_out,
                        float arc_i, float arc_j, float arc_k, float Q_w, float sigma, float dt,
                        float dx3, float eta, float pen, int enable, int nx, int ny, int nz)
{
    float inv2s2 = 1.0f / (2.0f * sigma * sigma);
    float w_total = or_arc_weight_sum(flags, f_l, arc_i, arc_j, arc_k, sigma, pen, enable, nx,
                                      ny, nz);
    if (norm_out) *norm_out = w_total;
    float epw = 0.0f;
    if (w_total > 1e-9f) epw = Q_w * eta * dt / (w_total * dx3);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (epw <= 0.0f || flags[c] == GAS) continue;
                float w_dep = arc_deposition_weight(k, arc_k, f_l[c], pen, enable);
                if (w_dep < 1e-6f) continue;
                float di = (float)i - arc_i, dj = (float)j - arc_j, dk = (float)k - arc_k;
                float r2 = di * di + dj * dj + dk * dk * 0.25f;
                H[c] += epw * waam_expf(-r2 * inv2s2) * w_dep;
            }
}

static inline float goldak_pdf_weight(float di, float dj, float dk, float sgn, float ff,
                                      float fr, float af, float ar, float b, float cax)
{
    const float eps = 1e-6f;
    float x = di * sgn;
    int front = x >= 0.0f;
    float a = front ? af : ar;
    float frac = front ? ff : fr;
    return frac * waam_expf(-3.0f * ((x * x) / (a * a + eps) + (dj * dj) / (b * b + eps) +
                                     (dk * dk) / (cax * cax + eps)));
}

void or_inject_goldak_heat(float *H, const uint8_t *flags, const float *f_l, float *norm_out,
                           float arc_i, float arc_j, float arc_k, float Q_w, float dt, float dx3,
                           float eta, float sgn, float ff, float fr, float af, float ar, float b,
                           float cax, float pen, int enable, int nx, int ny, int nz)
{
    csum_t acc = {0.0, 0.0};
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == GAS) continue;
                float w_dep = arc_deposition_weight(k, arc_k, f_l[c], pen, enable);
                if (w_dep < 1e-6f) continue;
                float w = goldak_pdf_weight((float)i - arc_i, (float)j - arc_j, (float)k - arc_k,
                                            sgn, ff, fr, af, ar, b, cax);
                csum_add(&acc, (double)(w * w_dep));
            }
    float w_total = (float)(acc.s + acc.c);
    if (norm_out) *norm_out = w_total;
    float epw = 0.0f;
    if (w_total > 1e-9f) epw = Q_w * eta * dt / (w_total * dx3);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (epw <= 0.0f || flags[c] == GAS) continue;
                float w_dep = arc_deposition_weight(k, arc_k, f_l[c], pen, enable);
                if (w_dep < 1e-6f) continue;
                float w = goldak_pdf_weight((float)i - arc_i, (float)j - arc_j, (float)k - arc_k,
                                            sgn, ff, fr, af, ar, b, cax);
                H[c] += epw * w * w_dep;
            }
}

/* surface_height_at kernels.py:625-641. The reference's scalar store races inside a
 * parallel range-for; the serial result (last qualifying k) is the maximum k. */
float or_surface_height_at(const float *phi, const uint8_t *flags, int i0, int j0,
                           int nz_fallback, int nx, int ny, int nz)
{
    (void)nx;
    int k_surf = imax(1, nz_fallback - 1);
    for (int k = 0; k < nz; ++k) {
        size_t c = IDX(i0, j0, k);
        if (flags[c] != GAS && phi[c] > 0.05f) k_surf = k;
    }
    return (float)k_surf;
}

/* ------------------------------------------------------------------------------------------
 * Wire deposition — _has_metal_neighbor kernels.py:245-275, feed_wire_surface :278-400.
 * Serial semantics: the three passes each sweep the sub-box in i, j, k order and the
 * reserve-then-commit budget is checked against the running fp32 accumulator.
 * ---------------------------------------------------------------------------------------- */
static inline int has_metal_neighbor(const float *phi, const uint8_t *flags, const float *f_l,
                                     int i, int j, int k, int nx, int ny, int nz)
{
    static const int D[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0},
                                {0, 1, 0},  {0, 0, -1}, {0, 0, 1}};
    int found = 0;
    for (int d = 0; d < 6; ++d) {
        int ii = i + D[d][0], jj = j + D[d][1], kk = k + D[d][2];
        if (ii >= 0 && jj >= 0 && kk >= 0 && ii < nx && jj < ny && kk < nz) {
            size_t c = IDX(ii, jj, kk);
            if (flags[c] == FLUID || flags[c] == SOLID) found = 1;
            else if (phi[c] > 0.35f || f_l[c] > 0.2f) found = 1;
        }
    }
    return found;
}

static inline void deposit_cell(float *f_src, uint8_t *flags, float *f_l, float *phi, float *H,
                                float *T, float *rho, size_t c, size_t N, float T_drop,
                                float cp_rho, float L_rho, float rho0)
{
    flags[c] = FLUID;
    f_l[c] = 1.0f;
    phi[c] = 1.0f;
    T[c] = T_drop;
    H[c] = cp_rho * T_drop + L_rho;
    rho[c] = rho0;
    for (int q = 0; q < Q; ++q) f_src[(size_t)q * N + c] = WQ[q] * rho0;
}

void or_feed_wire_surface(float *f_src, uint8_t *flags, float *f_l, float *phi, float *H,
                          float *T, float *rho, float arc_i, float arc_j, float arc_k,
                          float footprint_r, float droplet_radius, float target_vol,
                          float T_drop, float cp_rho, float L_rho, float rho0, float *vol_acc,
                          float *real_acc, float cell_vol, int nx, int ny, int nz)
{
    size_t N = NCELL;
    float search_r = fmaxf_(footprint_r, droplet_radius);
    int i_min = imax(0, (int)(arc_i - search_r));
    int i_max = imin(nx, (int)(arc_i + search_r) + 1);
    int j_min = imax(0, (int)(arc_j - search_r));
    int j_max = imin(ny, (int)(arc_j + search_r) + 1);
    int k_lo = imax(0, (int)(arc_k - 1.0f));
    float k_span = fmaxf_(droplet_radius + 3.0f, 4.0f);
    int k_hi = imin(nz, (int)(arc_k + k_span) + 1);
    float drop_cz = arc_k + droplet_radius + 0.5f;
    float r_drop2 = droplet_radius * droplet_radius;
    float foot2 = search_r * search_r;

    /* pass 1 (kernels.py:330-356) */
    for (int i = i_min; i < i_max; ++i)
        for (int j = j_min; j < j_max; ++j)
            for (int k = k_lo; k < k_hi; ++k) {
                size_t c = IDX(i, j, k);
                if (*vol_acc >= target_vol) continue;
                if (flags[c] != GAS) continue;
                if (!has_metal_neighbor(phi, flags, f_l, i, j, k, nx, ny, nz)) continue;
                float di = (float)i - arc_i, dj = (float)j - arc_j, dk = (float)k - drop_cz;
                float r2 = di * di + dj * dj + dk * dk;
                float r_ij2 = di * di + dj * dj;
                if (r_ij2 > foot2 && r2 > r_drop2) continue;
                float reserved = *vol_acc;
                *vol_acc = reserved + cell_vol;
                if (reserved >= target_vol) {
                    *vol_acc = *vol_acc - cell_vol;
                } else {
                    deposit_cell(f_src, flags, f_l, phi, H, T, rho, c, N, T_drop, cp_rho, L_rho,
                                 rho0);
                    *real_acc = *real_acc + cell_vol;
                }
            }
    /* pass 2 (kernels.py:359-383) */
    for (int i = i_min; i < i_max; ++i)
        for (int j = j_min; j < j_max; ++j)
            for (int k = k_lo; k < k_hi; ++k) {
                size_t c = IDX(i, j, k);
                if (*vol_acc >= target_vol) continue;
                if (flags[c] != GAS) continue;
                if (!has_metal_neighbor(phi, flags, f_l, i, j, k, nx, ny, nz)) continue;
                float di = (float)i - arc_i, dj = (float)j - arc_j;
                float r_ij2 = di * di + dj * dj;
                if (r_ij2 > foot2 * 1.44f) continue;
                float reserved = *vol_acc;
                *vol_acc = reserved + cell_vol;
                if (reserved >= target_vol) {
                    *vol_acc = *vol_acc - cell_vol;
                } else {
                    deposit_cell(f_src, flags, f_l, phi, H, T, rho, c, N, T_drop, cp_rho, L_rho,
                                 rho0);
                    *real_acc = *real_acc + cell_vol;
                }
            }
    /* pass 3 (kernels.py:388-400) */
    for (int i = i_min; i < i_max; ++i)
        for (int j = j_min; j < j_max; ++j)
            for (int k = k_lo; k < k_hi; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] != FLUID || f_l[c] < 0.45f) continue;
                float di = (float)i - arc_i, dj = (float)j - arc_j;
                float r_ij2 = di * di + dj * dj;
                if (r_ij2 > foot2 * 1.44f) continue;
                T[c] = T_drop;
                H[c] = cp_rho * T_drop + L_rho;
                rho[c] = rho0;
                for (int q = 0; q < Q; ++q) f_src[(size_t)q * N + c] = WQ[q] * rho0;
            }
}

/* feed_wire_momentum_impact kernels.py:1176-1233 */
void or_feed_wire_momentum_impact(float *f_src, const float *rho, float *ux, float *uy,
                                  float *uz, const uint8_t *flags, const float *f_l, float arc_i,
                                  float arc_j, float arc_k, float dr, float vx, float vy,
                                  float vz, int nx, int ny, int nz)
{
    size_t N = NCELL;
    int i_min = imax(0, (int)(arc_i - dr * 2.5f));
    int i_max = imin(nx, (int)(arc_i + dr * 2.5f) + 1);
    int j_min = imax(0, (int)(arc_j - dr * 2.5f));
    int j_max = imin(ny, (int)(arc_j + dr * 2.5f) + 1);
    int k_min = imax(0, (int)(arc_k));
    int k_max = imin(nz, (int)(arc_k + dr * 3.0f) + 1);
    float drop_cz = arc_k + dr + 1.0f;
    float u2 = vx * vx + vy * vy + vz * vz;
    for (int i = i_min; i < i_max; ++i)
        for (int j = j_min; j < j_max; ++j)
            for (int k = k_min; k < k_max; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == GAS) continue;
                if (f_l[c] < 0.45f) continue;
                float di = (float)i - arc_i, dj = (float)j - arc_j, dk = (float)k - drop_cz;
                float r2 = di * di + dj * dj + dk * dk;
                float r_ij2 = di * di + dj * dj;
                int in_drop = r2 <= dr * dr || (r_ij2 <= dr * dr && (float)k >= arc_k + 1.0f);
                if (in_drop) {
                    float r_loc = rho[c];
                    ux[c] = vx;
                    uy[c] = vy;
                    uz[c] = vz;
                    for (int q = 0; q < Q; ++q) {
                        float eu = (float)EX[q] * vx + (float)EY[q] * vy + (float)EZ[q] * vz;
                        f_src[(size_t)q * N + c] =
                            WQ[q] * r_loc * (1.0f + 3.0f * eu + 4.5f * eu * eu - 1.5f * u2);
                    }
                }
            }
}

/* apply_droplet_impact_pressure kernels.py:1142-1173 (+ _wf_pressure_to_Fz_lu :1033-1035) */
void or_apply_droplet_impact_pressure(float *Fz, const uint8_t *flags, const float *phi,
                                      const float *f_l, float arc_i, float arc_j, float arc_k,
                                      float dr, float impact_pa, float dt, float dx,
                                      float rho_ref, int nx, int ny, int nz)
{
    const float eps = 1e-6f;
    float inv2s2 = 1.0f / (2.0f * dr * dr + eps);
    float F_peak_phys = impact_pa / (rho_ref * dx);
    float F_peak_lu = F_peak_phys * dt * dt / dx;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == GAS) continue;
                if (f_l[c] < 0.25f && phi[c] < 0.25f) continue;
                float di = (float)i - arc_i, dj = (float)j - arc_j, dk = (float)k - arc_k;
                float r2 = di * di + dj * dj + dk * dk;
                if (r2 > dr * dr * 4.0f) continue;
                Fz[c] += -F_peak_lu * waam_expf(-r2 * inv2s2);
            }
}

/* ------------------------------------------------------------------------------------------
 * CSF surface tension — _solid_wall_normal kernels.py:680-719,
 * _correct_normal_contact_angle :722-766, _phi_unit_normal_at :769-796,
 * _brackbill_curvature_at :799-855, compute_csf_tension :858-926.
 * ---------------------------------------------------------------------------------------- */
static inline void phi_unit_normal_at(const float *phi, int i, int j, int k, int nx, int ny,
                                      int nz, float eps, float *onx, float *ony, float *onz,
                                      float *og)
{
    float dpx = 0.5f * (phi[IDX(imin(i + 1, nx - 1), j, k)] - phi[IDX(imax(i - 1, 0), j, k)]);
    float dpy = 0.5f * (phi[IDX(i, imin(j + 1, ny - 1), k)] - phi[IDX(i, imax(j - 1, 0), k)]);
    float dpz = 0.5f * (phi[IDX(i, j, imin(k + 1, nz - 1))] - phi[IDX(i, j, imax(k - 1, 0))]);
    float g = sqrtf(dpx * dpx + dpy * dpy + dpz * dpz);
    float a = 0.0f, b = 0.0f, c = 0.0f;
    if (g >= eps) {
        a = dpx / g;
        b = dpy / g;
        c = dpz / g;
    }
    *onx = a;
    *ony = b;
    *onz = c;
    *og = g;
}

static inline void brackbill_curvature_at(const float *phi, int i, int j, int k, int nx, int ny,
                                          int nz, float eps, float *okappa, float *og0)
{
    float nx0, ny0, nz0, g0;
    phi_unit_normal_at(phi, i, j, k, nx, ny, nz, eps, &nx0, &ny0, &nz0, &g0);
    float kappa = 0.0f;
    if (g0 >= eps) {
        float t1, t2, nxp, nxm, nyp, nym, nzp, nzm, gp, gm, gp2, gm2, gp3, gm3;
        phi_unit_normal_at(phi, imin(i + 1, nx - 1), j, k, nx, ny, nz, eps, &nxp, &t1, &t2, &gp);
        phi_unit_normal_at(phi, imax(i - 1, 0), j, k, nx, ny, nz, eps, &nxm, &t1, &t2, &gm);
        phi_unit_normal_at(phi, i, imin(j + 1, ny - 1), k, nx, ny, nz, eps, &t1, &nyp, &t2, &gp2);
        phi_unit_normal_at(phi, i, imax(j - 1, 0), k, nx, ny, nz, eps, &t1, &nym, &t2, &gm2);
        phi_unit_normal_at(phi, i, j, imin(k + 1, nz - 1), nx, ny, nz, eps, &t1, &t2, &nzp, &gp3);
        phi_unit_normal_at(phi, i, j, imax(k - 1, 0), nx, ny, nz, eps, &t1, &t2, &nzm, &gm3);
        if (gp < eps) nxp = nx0;
        if (gm < eps) nxm = nx0;
        if (gp2 < eps) nyp = ny0;
        if (gm2 < eps) nym = ny0;
        if (gp3 < eps) nzp = nz0;
        if (gm3 < eps) nzm = nz0;
        float div_n = 0.5f * (nxp - nxm) + 0.5f * (nyp - nym) + 0.5f * (nzp - nzm);
        kappa = -div_n;
    }
    *okappa = kappa;
    *og0 = g0;
}

static inline int solid_wall_normal(const uint8_t *flags, int i, int j, int k, int nx, int ny,
                                    int nz, float *owx, float *owy, float *owz)
{
    float wx = 0.0f, wy = 0.0f, wz = 0.0f;
    int has_wall = 0;
    if (i > 0 && flags[IDX(i - 1, j, k)] == SOLID) { wx += 1.0f; has_wall = 1; }
    if (i < nx - 1 && flags[IDX(i + 1, j, k)] == SOLID) { wx -= 1.0f; has_wall = 1; }
    if (j > 0 && flags[IDX(i, j - 1, k)] == SOLID) { wy += 1.0f; has_wall = 1; }
    if (j < ny - 1 && flags[IDX(i, j + 1, k)] == SOLID) { wy -= 1.0f; has_wall = 1; }
    if (k > 0 && flags[IDX(i, j, k - 1)] == SOLID) { wz += 1.0f; has_wall = 1; }
    if (k < nz - 1 && flags[IDX(i, j, k + 1)] == SOLID) { wz -= 1.0f; has_wall = 1; }
    float wmag = sqrtf(wx * wx + wy * wy + wz * wz);
    if (wmag > 1e-8f) {
        wx /= wmag;
        wy /= wmag;
        wz /= wmag;
    }
    *owx = wx;
    *owy = wy;
    *owz = wz;
    return has_wall;
}

static inline void correct_normal_contact_angle(float nxn, float nyn, float nzn, float wx,
                                                float wy, float wz, float cos_t, float sin_t,
                                                float *ox, float *oy, float *oz)
{
    float dot = nxn * wx + nyn * wy + nzn * wz;
    float tx = nxn - dot * wx, ty = nyn - dot * wy, tz = nzn - dot * wz;
    float tmag = sqrtf(tx * tx + ty * ty + tz * tz);
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    if (tmag > 1e-8f) {
        cx = cos_t * wx + sin_t * tx / tmag;
        cy = cos_t * wy + sin_t * ty / tmag;
        cz = cos_t * wz + sin_t * tz / tmag;
    } else {
        if (fabsf(wz) > 0.5f) {
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
    }
    float cmag = sqrtf(cx * cx + cy * cy + cz * cz);
    if (cmag > 1e-8f) {
        cx /= cmag;
        cy /= cmag;
        cz /= cmag;
    }
    *ox = cx;
    *oy = cy;
    *oz = cz;
}

void or_compute_csf_tension(const float *phi, const uint8_t *flags, float *Fx, float *Fy,
                            float *Fz, float gamma_lu, int enable_wetting, float cos_t,
                            float sin_t, int nx, int ny, int nz)
{
    const float eps = 1e-6f;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == SOLID || flags[c] == GAS) continue;
                if (i < 1 || j < 1 || k < 1 || i > nx - 2 || j > ny - 2 || k > nz - 2) continue;
                float kappa, gmag;
                brackbill_curvature_at(phi, i, j, k, nx, ny, nz, eps, &kappa, &gmag);
                if (gmag < eps) continue;
                float dpx = 0.5f * (phi[IDX(imin(i + 1, nx - 1), j, k)] - phi[IDX(imax(i - 1, 0), j, k)]);
                float dpy = 0.5f * (phi[IDX(i, imin(j + 1, ny - 1), k)] - phi[IDX(i, imax(j - 1, 0), k)]);
                float dpz = 0.5f * (phi[IDX(i, j, imin(k + 1, nz - 1))] - phi[IDX(i, j, imax(k - 1, 0))]);
                float fx = gamma_lu * kappa * dpx;
                float fy = gamma_lu * kappa * dpy;
                float fz = gamma_lu * kappa * dpz;
                if (enable_wetting != 0) {
                    float wx, wy, wz;
                    int has_wall = solid_wall_normal(flags, i, j, k, nx, ny, nz, &wx, &wy, &wz);
                    if (has_wall != 0) {
                        float nxn = dpx / gmag, nyn = dpy / gmag, nzn = dpz / gmag;
                        float cx, cy, cz;
                        correct_normal_contact_angle(nxn, nyn, nzn, wx, wy, wz, cos_t, sin_t,
                                                     &cx, &cy, &cz);
                        fx = gamma_lu * kappa * cx * gmag;
                        fy = gamma_lu * kappa * cy * gmag;
                        fz = gamma_lu * kappa * cz * gmag;
                    }
                }
                Fx[c] += fx;
                Fy[c] += fy;
                Fz[c] += fz;
            }
}

/* compute_curvature_field kernels.py:3353-3376 (export/validation helper) */
void or_compute_curvature_field(const float *phi, const uint8_t *flags, float *kappa_out, int nx,
                                int ny, int nz)
{
    const float eps = 1e-6f;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                kappa_out[c] = 0.0f;
                if (flags[c] == GAS) continue;
                if (i < 1 || j < 1 || k < 1 || i > nx - 2 || j > ny - 2 || k > nz - 2) continue;
                float kappa, g;
                brackbill_curvature_at(phi, i, j, k, nx, ny, nz, eps, &kappa, &g);
                if (g >= eps) kappa_out[c] = kappa;
            }
}

/* ------------------------------------------------------------------------------------------
 * Contact-angle ghost-phi BC — apply_contact_angle_phi_bc kernels.py:933-991.
 * serial: in-place lexicographic (ti.cpu single thread).
 * jacobi: neighbour reads from the pre-kernel phi; the one cross-cell write of the
 * reference (rule 2: cell k raises phi at k+1) is gathered by cell k+1 and applied before
 * that cell's own rules, as serial order applies it.
 * ---------------------------------------------------------------------------------------- */
void or_apply_contact_angle_phi_bc_serial(float *phi, const uint8_t *flags, float cos_t,
                                          float sin_t, int nx, int ny, int nz)
{
    float film = 0.5f * (1.0f - cos_t);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == SOLID) continue;
                int below_solid = (k > 0 && flags[IDX(i, j, k - 1)] == SOLID);
                if (below_solid && flags[c] == GAS) {
                    if (phi[c] < film) phi[c] = film;
                }
                if (below_solid && phi[c] > 0.55f) {
                    if (k + 1 < nz && flags[IDX(i, j, k + 1)] == GAS) {
                        float g = phi[c] - sin_t;
                        g = fmaxf_(film, fminf_(g, 0.98f));
                        if (phi[IDX(i, j, k + 1)] < g) phi[IDX(i, j, k + 1)] = g;
                    }
                }
                if (k > 0 && phi[IDX(i, j, k - 1)] > 0.55f && flags[c] == GAS) {
                    float t = phi[IDX(i, j, k - 1)] - sin_t;
                    t = fmaxf_(film, fminf_(t, 0.95f));
                    if (phi[c] < t) phi[c] = t;
                }
                if (phi[c] < 0.45f) {
                    for (int d = -1; d <= 1; d += 2) {
                        int jj = j + d;
                        if (jj >= 0 && jj < ny && phi[IDX(i, jj, k)] > 0.55f) {
                            float spread = film * (0.85f + 0.15f * sin_t);
                            if (phi[c] < spread) phi[c] = spread;
                        }
                    }
                    for (int d = -1; d <= 1; d += 2) {
                        int ii = i + d;
                        if (ii >= 0 && ii < nx && phi[IDX(ii, j, k)] > 0.55f) {
                            float spread = film * (0.85f + 0.15f * sin_t);
                            if (phi[c] < spread) phi[c] = spread;
                        }
                    }
                }
                if (below_solid && phi[c] > 0.2f) {
                    if (phi[c] < 0.65f) phi[c] = 0.65f;
                }
            }
}

void or_apply_contact_angle_phi_bc_jacobi(const float *phi_in, float *phi_out,
                                          const uint8_t *flags, float cos_t, float sin_t, int nx,
                                          int ny, int nz)
{
    float film = 0.5f * (1.0f - cos_t);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                float p = phi_in[c];
                if (flags[c] == SOLID) {
                    phi_out[c] = p;
                    continue;
                }
                /* rule 2 of the cell below, aimed at this cell */
                if (k >= 2 && flags[c] == GAS && flags[IDX(i, j, k - 2)] == SOLID &&
                    flags[IDX(i, j, k - 1)] != SOLID) {
                    float pb = phi_in[IDX(i, j, k - 1)];
                    if (flags[IDX(i, j, k - 1)] == GAS && pb < film) pb = film; /* its rule 1 */
                    if (pb > 0.55f) {
                        float g = pb - sin_t;
                        g = fmaxf_(film, fminf_(g, 0.98f));
                        if (p < g) p = g;
                    }
                }
                int below_solid = (k > 0 && flags[IDX(i, j, k - 1)] == SOLID);
                if (below_solid && flags[c] == GAS) {
                    if (p < film) p = film;
                }
                if (k > 0 && phi_in[IDX(i, j, k - 1)] > 0.55f && flags[c] == GAS) {
                    float t = phi_in[IDX(i, j, k - 1)] - sin_t;
                    t = fmaxf_(film, fminf_(t, 0.95f));
                    if (p < t) p = t;
                }
                if (p < 0.45f) {
                    for (int d = -1; d <= 1; d += 2) {
                        int jj = j + d;
                        if (jj >= 0 && jj < ny && phi_in[IDX(i, jj, k)] > 0.55f) {
                            float spread = film * (0.85f + 0.15f * sin_t);
                            if (p < spread) p = spread;
                        }
                    }
                    for (int d = -1; d <= 1; d += 2) {
                        int ii = i + d;
                        if (ii >= 0 && ii < nx && phi_in[IDX(ii, j, k)] > 0.55f) {
                            float spread = film * (0.85f + 0.15f * sin_t);
                            if (p < spread) p = spread;
                        }
                    }
                }
                if (below_solid && p > 0.2f) {
                    if (p < 0.65f) p = 0.65f;
                }
                phi_out[c] = p;
            }
}

/* ------------------------------------------------------------------------------------------
 * Thermal — update_phase kernels.py:1519-1556; _minmod :1565-1571; _masked_at :1574-1590;
 * _limited_upwind_grad :1593-1608; advect_diffuse_temperature :1611-1682;
 * apply_thermal_boundary_losses :2030-2082; clamp_enthalpy_floor_scalar :1937-1950;
 * clamp_enthalpy_ceiling_scalar :1953-1984; update_cooling_rate :2013-2027;
 * update_T_max :3016-3032; update_time_above_T :3261-3283.
 * ---------------------------------------------------------------------------------------- */
void or_update_phase(const float *H, float *T, float *f_l, float cp_rho, float L_rho,
                     float T_s, float T_l, int nx, int ny, int nz)
{
    float H_sol = cp_rho * T_s;
    float H_liq = H_sol + L_rho;
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        float h = H[c];
        if (h <= H_sol) {
            f_l[c] = 0.0f;
            T[c] = h / cp_rho;
        } else if (h >= H_liq) {
            f_l[c] = 1.0f;
            T[c] = T_l + (h - H_liq) / cp_rho;
        } else {
            float fl = (h - H_sol) / L_rho;
            f_l[c] = fl;
            T[c] = T_s + fl * (T_l - T_s);
        }
    }
}

static inline float minmod(float a, float b)
{
    float out = 0.0f;
    if (a * b > 0.0f) out = (fabsf(a) < fabsf(b)) ? a : b;
    return out;
}

static inline float masked_at(const float *fld, const uint8_t *flags, int i, int j, int k,
                              float fallback, int nx, int ny, int nz)
{
    int ii = imax(0, imin(i, nx - 1));
    int jj = imax(0, imin(j, ny - 1));
    int kk = imax(0, imin(k, nz - 1));
    size_t c = IDX(ii, jj, kk);
    float v = fld[c];
    if (flags[c] == GAS) v = fallback;
    return v;
}

static inline float limited_upwind_grad(float u, float qmm, float qm, float qc, float qp,
                                        float qpp)
{
    float g;
    if (u > 0.0f) {
        float d1 = qc - qm;
        float d2 = 0.5f * (3.0f * qc - 4.0f * qm + qmm);
        g = minmod(d1, d2);
    } else {
        float d1 = qp - qc;
        float d2 = 0.5f * (-3.0f * qc + 4.0f * qp - qpp);
        g = minmod(d1, d2);
    }
    return g;
}

/* dH for one cell; Hs is the field the H stencil reads (in-place H for serial, the pre-kernel
 * copy for Jacobi). alpha/cp_rho are per-call scalars or per-cell values (variable form). */
static inline float thermal_dH(const float *Hs, const float *T, const float *ux, const float *uy,
                               const float *uz, const uint8_t *flags, int i, int j, int k,
                               float alpha_lu, float dt, float cp_rho, int nx, int ny, int nz)
{
    size_t c = IDX(i, j, k);
    uint8_t flag = flags[c];
    float T_c = T[c];
    float H_c = Hs[c];
    float T_im = masked_at(T, flags, i - 1, j, k, T_c, nx, ny, nz);
    float T_ip = masked_at(T, flags, i + 1, j, k, T_c, nx, ny, nz);
    float T_jm = masked_at(T, flags, i, j - 1, k, T_c, nx, ny, nz);
    float T_jp = masked_at(T, flags, i, j + 1, k, T_c, nx, ny, nz);
    float T_km = masked_at(T, flags, i, j, k - 1, T_c, nx, ny, nz);
    float T_kp = masked_at(T, flags, i, j, k + 1, T_c, nx, ny, nz);
    float lap = T_ip + T_im + T_jp + T_jm + T_kp + T_km - 6.0f * T_c;
    float dH = cp_rho * alpha_lu * lap * dt;
    if (flag != SOLID) {
        float u = ux[c], v = uy[c], w = uz[c];
        float H_im = masked_at(Hs, flags, i - 1, j, k, H_c, nx, ny, nz);
        float H_ip = masked_at(Hs, flags, i + 1, j, k, H_c, nx, ny, nz);
        float H_imm = masked_at(Hs, flags, i - 2, j, k, H_im, nx, ny, nz);
        float H_ipp = masked_at(Hs, flags, i + 2, j, k, H_ip, nx, ny, nz);
        float H_jm = masked_at(Hs, flags, i, j - 1, k, H_c, nx, ny, nz);
        float H_jp = masked_at(Hs, flags, i, j + 1, k, H_c, nx, ny, nz);
        float H_jmm = masked_at(Hs, flags, i, j - 2, k, H_jm, nx, ny, nz);
        float H_jpp = masked_at(Hs, flags, i, j + 2, k, H_jp, nx, ny, nz);
        float H_km = masked_at(Hs, flags, i, j, k - 1, H_c, nx, ny, nz);
        float H_kp = masked_at(Hs, flags, i, j, k + 1, H_c, nx, ny, nz);
        float H_kmm = masked_at(Hs, flags, i, j, k - 2, H_km, nx, ny, nz);
        float H_kpp = masked_at(Hs, flags, i, j, k + 2, H_kp, nx, ny, nz);
        float gx = limited_upwind_grad(u, H_imm, H_im, H_c, H_ip, H_ipp);
        float gy = limited_upwind_grad(v, H_jmm, H_jm, H_c, H_jp, H_jpp);
        float gz = limited_upwind_grad(w, H_kmm, H_km, H_c, H_kp, H_kpp);
        dH -= (u * gx + v * gy + w * gz) * dt;
    }
    return dH;
}

void or_advect_diffuse_temperature_serial(float *H, const float *T, const float *ux,
                                          const float *uy, const float *uz,
                                          const uint8_t *flags, float alpha_lu, float dt,
                                          float cp_rho, int nx, int ny, int nz)
{
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == GAS) continue;
                float dH = thermal_dH(H, T, ux, uy, uz, flags, i, j, k, alpha_lu, dt, cp_rho, nx,
                                      ny, nz);
                H[c] += dH;
            }
}

void or_advect_diffuse_temperature_jacobi(const float *H_in, float *H_out, const float *T,
                                          const float *ux, const float *uy, const float *uz,
                                          const uint8_t *flags, float alpha_lu, float dt,
                                          float cp_rho, int nx, int ny, int nz)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == GAS) {
                    H_out[c] = H_in[c];
                    continue;
                }
                float dH = thermal_dH(H_in, T, ux, uy, uz, flags, i, j, k, alpha_lu, dt, cp_rho,
                                      nx, ny, nz);
                H_out[c] = H_in[c] + dH;
            }
}

void or_apply_thermal_boundary_losses(float *H, const float *T, const uint8_t *flags,
                                      float T_amb, float h_conv, float eps_rad, int enable_conv,
                                      int enable_rad, float cp_rho, float dt, float dx,
                                      float sigma_sb, int nx, int ny, int nz)
{
    (void)cp_rho;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == GAS) continue;
                int n_exp = 0;
                if (k + 1 < nz && flags[IDX(i, j, k + 1)] == GAS) n_exp += 1;
                if (j + 1 < ny && flags[IDX(i, j + 1, k)] == GAS) n_exp += 1;
                if (j - 1 >= 0 && flags[IDX(i, j - 1, k)] == GAS) n_exp += 1;
                if (i + 1 < nx && flags[IDX(i + 1, j, k)] == GAS) n_exp += 1;
                if (i - 1 >= 0 && flags[IDX(i - 1, j, k)] == GAS) n_exp += 1;
                if (n_exp == 0) continue;
                float T_eff = fminf_(T[c], 2500.0f);
                float q_loss = 0.0f;
                if (enable_conv == 1) q_loss += h_conv * (T_eff - T_amb);
                if (enable_rad == 1) {
                    float t2 = T_eff * T_eff, a2 = T_amb * T_amb;
                    q_loss += eps_rad * sigma_sb * (t2 * t2 - a2 * a2);
                }
                H[c] -= q_loss * (float)n_exp * dt / dx;
            }
}

void or_clamp_enthalpy_floor_scalar(float *H, float cp_rho, const uint8_t *flags, float T_floor,
                                    int nx, int ny, int nz)
{
    float h_min = cp_rho * T_floor;
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] == GAS) continue;
        if (H[c] < h_min) H[c] = h_min;
    }
}

void or_clamp_enthalpy_ceiling_scalar(float *H, const uint8_t *flags, float cp_rho, float T_s,
                                      float T_l, float T_cap, float L_rho, int nx, int ny, int nz)
{
    float H_sol = cp_rho * T_s;
    float H_liq = H_sol + L_rho;
    float h_cap_liquid = H_liq + cp_rho * (T_cap - T_l);
    float h_cap_solid = cp_rho * T_cap;
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] == GAS) continue;
        float h = H[c];
        if (h > H_liq) {
            if (h > h_cap_liquid) H[c] = h_cap_liquid;
        } else {
            if (h > h_cap_solid) H[c] = h_cap_solid;
        }
    }
}

void or_update_cooling_rate(const float *T, float *T_prev, float *dT_dt, const uint8_t *flags,
                            float dt_phys, int nx, int ny, int nz)
{
    float inv_dt = 1.0f / (dt_phys + 1e-12f);
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] == GAS) continue;
        dT_dt[c] = (T[c] - T_prev[c]) * inv_dt;
        T_prev[c] = T[c];
    }
}

void or_update_T_max(const float *T, float *T_max, const uint8_t *flags, int nx, int ny, int nz)
{
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] == GAS) continue;
        if (T[c] > T_max[c]) T_max[c] = T[c];
    }
}

void or_update_time_above_T(const float *T, const uint8_t *flags, float *t800, float *t1100,
                            float *tsol, float dt, float T_800, float T_1100, float T_s, int nx,
                            int ny, int nz)
{
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] == GAS) continue;
        float t = T[c];
        if (t >= T_800) t800[c] += dt;
        if (t >= T_1100) t1100[c] += dt;
        if (t >= T_s) tsol[c] += dt;
    }
}

/* ------------------------------------------------------------------------------------------
 * Forces — compute_marangoni_force kernels.py:2158-2217; add_buoyancy :2729-2753;
 * add_hydrostatic_gravity :2756-2772; apply_arc_pressure :3038-3083;
 * clamp_body_force_magnitude :3286-3309.
 * ---------------------------------------------------------------------------------------- */
void or_compute_marangoni_force(const float *T, const float *phi, const float *f_l, float *Fx,
                                float *Fy, float *Fz, const uint8_t *flags, float dgamma,
                                int nx, int ny, int nz)
{
    const float eps = 1e-6f, fl_min = 0.05f;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == SOLID || flags[c] == GAS) continue;
                if (f_l[c] < fl_min) continue;
                int ip = imin(i + 1, nx - 1), im = imax(i - 1, 0);
                int jp = imin(j + 1, ny - 1), jm = imax(j - 1, 0);
                int kp = imin(k + 1, nz - 1), km = imax(k - 1, 0);
                float dphx = 0.5f * (phi[IDX(ip, j, k)] - phi[IDX(im, j, k)]);
                float dphy = 0.5f * (phi[IDX(i, jp, k)] - phi[IDX(i, jm, k)]);
                float dphz = 0.5f * (phi[IDX(i, j, kp)] - phi[IDX(i, j, km)]);
                float gm = sqrtf(dphx * dphx + dphy * dphy + dphz * dphz);
                if (gm < eps) continue;
                float nxn = dphx / gm, nyn = dphy / gm, nzn = dphz / gm;
                float dTx = 0.5f * (T[IDX(ip, j, k)] - T[IDX(im, j, k)]);
                float dTy = 0.5f * (T[IDX(i, jp, k)] - T[IDX(i, jm, k)]);
                float dTz = 0.5f * (T[IDX(i, j, kp)] - T[IDX(i, j, km)]);
                float ndot = nxn * dTx + nyn * dTy + nzn * dTz;
                float sx = dTx - ndot * nxn, sy = dTy - ndot * nyn, sz = dTz - ndot * nzn;
                float scale = dgamma * gm;
                Fx[c] += scale * sx;
                Fy[c] += scale * sy;
                Fz[c] += scale * sz;
            }
}

void or_add_buoyancy(const float *T, float *Fz, const float *f_l, const uint8_t *flags,
                     float g_lu, float beta, float T_ref, float rho_ref, int nx, int ny, int nz)
{
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] == SOLID || flags[c] == GAS) continue;
        float fl = f_l[c];
        if (fl > 0.0f) Fz[c] += rho_ref * g_lu * beta * (T[c] - T_ref) * fl;
    }
}

void or_add_hydrostatic_gravity(float *Fz, const float *f_l, const uint8_t *flags,
                                float rho_ref, float g_lu, int nx, int ny, int nz)
{
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] == SOLID || flags[c] == GAS) continue;
        float fl = f_l[c];
        if (fl > 0.0f) Fz[c] += -rho_ref * g_lu * fl;
    }
}

void or_apply_arc_pressure(float *Fz, const uint8_t *flags, const float *phi, float arc_i,
                           float arc_j, float arc_k, float sigma, float pressure_pa, float dt,
                           float dx, float rho_ref, int nx, int ny, int nz)
{
    const float eps = 1e-6f;
    float inv2s2 = 1.0f / (2.0f * sigma * sigma + eps);
    float F_peak_phys = pressure_pa / (rho_ref * dx);
    float F_peak_lu = F_peak_phys * (dt * dt) / dx;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == SOLID || flags[c] == GAS) continue;
                float dphz = phi[IDX(i, j, imin(k + 1, nz - 1))] - phi[IDX(i, j, imax(k - 1, 0))];
                if (fabsf(dphz) > 0.05f) {
                    float di = (float)i - arc_i, dj = (float)j - arc_j, dk = (float)k - arc_k;
                    float r2 = di * di + dj * dj + dk * dk * 0.1f;
                    float force = -F_peak_lu * waam_expf(-r2 * inv2s2);
                    Fz[c] += force;
                }
            }
}

void or_clamp_body_force_magnitude(float *Fx, float *Fy, float *Fz, const uint8_t *flags,
                                   float F_max, int nx, int ny, int nz)
{
    const float eps = 1e-12f;
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] == SOLID || flags[c] == GAS) continue;
        float fx = Fx[c], fy = Fy[c], fz = Fz[c];
        float mag = sqrtf(fx * fx + fy * fy + fz * fz);
        if (mag > F_max) {
            float s = F_max / (mag + eps);
            Fx[c] = fx * s;
            Fy[c] = fy * s;
            Fz[c] = fz * s;
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * VOF — _vof_face_flux kernels.py:2275-2317; advect_phi :2320-2361; reinitialize_phi
 * :2364-2391; update_flags_from_phi :2394-2419; sync_phi_liquid_fraction :3484-3504.
 * ---------------------------------------------------------------------------------------- */
static inline float vof_face_flux(const float *phi, const float *vel, const uint8_t *flags,
                                  int i0, int j0, int k0, int i1, int j1, int k1, int nx, int ny,
                                  int nz)
{
    float flux = 0.0f;
    int ii0 = (i0 + nx) % nx, ii1 = (i1 + nx) % nx;
    int in0 = j0 >= 0 && k0 >= 0 && j0 < ny && k0 < nz;
    int in1 = j1 >= 0 && k1 >= 0 && j1 < ny && k1 < nz;
    if (in0 && in1) {
        size_t c0 = IDX(ii0, j0, k0), c1 = IDX(ii1, j1, k1);
        uint8_t f0 = flags[c0], f1 = flags[c1];
        if (f0 != SOLID && f1 != SOLID) {
            float u0 = 0.0f, u1 = 0.0f;
            if (f0 != GAS) u0 = vel[c0];
            if (f1 != GAS) u1 = vel[c1];
            float uf = 0.5f * (u0 + u1);
            if (uf > 0.0f) flux = uf * phi[c0];
            else flux = uf * phi[c1];
        }
    }
    return flux;
}

void or_advect_phi(float *phi_dst, const float *phi_src, const float *ux, const float *uy,
                   const float *uz, const uint8_t *flags, int nx, int ny, int nz)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == SOLID) {
                    phi_dst[c] = phi_src[c];
                    continue;
                }
                float fe = vof_face_flux(phi_src, ux, flags, i, j, k, i + 1, j, k, nx, ny, nz);
                float fw = vof_face_flux(phi_src, ux, flags, i - 1, j, k, i, j, k, nx, ny, nz);
                float fn = vof_face_flux(phi_src, uy, flags, i, j, k, i, j + 1, k, nx, ny, nz);
                float fs = vof_face_flux(phi_src, uy, flags, i, j - 1, k, i, j, k, nx, ny, nz);
                float ft = vof_face_flux(phi_src, uz, flags, i, j, k, i, j, k + 1, nx, ny, nz);
                float fb = vof_face_flux(phi_src, uz, flags, i, j, k - 1, i, j, k, nx, ny, nz);
                phi_dst[c] = phi_src[c] - (fe - fw) - (fn - fs) - (ft - fb);
            }
}

void or_reinitialize_phi(float *phi, const uint8_t *flags, int nx, int ny, int nz)
{
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        uint8_t f = flags[c];
        if (f == SOLID) phi[c] = 1.0f;
        else if (f == GAS) {
            if (phi[c] < 1e-6f) phi[c] = 0.0f;
            else phi[c] = fminf_(1.0f, phi[c]);
        } else {
            float p = phi[c];
            phi[c] = fmaxf_(0.0f, fminf_(1.0f, p));
        }
    }
}

void or_update_flags_from_phi(const float *phi, const float *f_l, uint8_t *flags, int nz_solid,
                              int nx, int ny, int nz)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (k < nz_solid) {
                    flags[c] = SOLID;
                    continue;
                }
                float p = phi[c], fl = f_l[c];
                if (p > 0.85f && fl < 0.05f) flags[c] = SOLID;
                else if (p > 0.55f && fl > 0.1f) flags[c] = FLUID;
                else if (p < 0.05f) flags[c] = GAS;
                else flags[c] = IFACE;
            }
}

void or_sync_phi_liquid_fraction(const float *phi, float *f_l, const uint8_t *flags, int nx,
                                 int ny, int nz)
{
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        uint8_t f = flags[c];
        if (f == GAS || phi[c] < 0.05f) f_l[c] = 0.0f;
        else if (f == SOLID) f_l[c] = 0.0f;
    }
}

/* ------------------------------------------------------------------------------------------
 * Freeze / remelt — solidify_cooled_metal kernels.py:2422-2450; remelt_hot_solid_scalar
 * :2489-2520; solidify_trailing_pool_scalar :2573-2617. Pointwise.
 * ---------------------------------------------------------------------------------------- */
void or_solidify_cooled_metal(const float *T, float *f_l, float *phi, uint8_t *flags, float *ux,
                              float *uy, float *uz, float T_s, int zero_velocity, int nx, int ny,
                              int nz)
{
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] == GAS) continue;
        if (phi[c] < 0.85f) continue;
        if (f_l[c] < 0.02f && T[c] < T_s) {
            flags[c] = SOLID;
            phi[c] = 1.0f;
            f_l[c] = 0.0f;
            if (zero_velocity == 1) {
                ux[c] = 0.0f;
                uy[c] = 0.0f;
                uz[c] = 0.0f;
            }
        }
    }
}

void or_remelt_hot_solid_scalar(const float *T, float *H, float *f_l, float *phi,
                                uint8_t *flags, float cp_rho, float T_s, float T_l, float L_rho,
                                int nx, int ny, int nz)
{
    const float margin = 5.0f;
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] != SOLID) continue;
        float T_c = T[c];
        if (T_c <= T_s + margin) continue;
        float H_sol = cp_rho * T_s;
        float fl;
        if (T_c >= T_l) {
            fl = 1.0f;
            H[c] = cp_rho * T_l + L_rho + cp_rho * (T_c - T_l);
        } else {
            fl = (T_c - T_s) / (T_l - T_s + 1e-6f);
            H[c] = H_sol + fl * L_rho;
        }
        f_l[c] = fl;
        phi[c] = 1.0f;
        flags[c] = FLUID;
    }
}

void or_solidify_trailing_pool_scalar(float *T, float *H, float *f_l, float *phi, uint8_t *flags,
                                      float *ux, float *uy, float *uz, float cp_rho, float arc_i,
                                      float arc_j, float arc_k, float dir_x, float dir_y,
                                      float dir_z, float lookback, float T_freeze, float T_s,
                                      int nx, int ny, int nz)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == GAS) continue;
                if (phi[c] < 0.85f || f_l[c] < 0.05f) continue;
                float di = (float)i - arc_i, dj = (float)j - arc_j, dk = (float)k - arc_k;
                float proj = di * dir_x + dj * dir_y + dk * dir_z;
                if (proj > -lookback) continue;
                if (T[c] >= T_freeze) continue;
                H[c] = cp_rho * (T_s - 1.0f);
                T[c] = T_s - 1.0f;
                f_l[c] = 0.0f;
                phi[c] = 1.0f;
                flags[c] = SOLID;
                ux[c] = 0.0f;
                uy[c] = 0.0f;
                uz[c] = 0.0f;
            }
}

/* ------------------------------------------------------------------------------------------
 * LBM — collide_srt kernels.py:2778-2885; collide_srt_variable_tau :2888-2965;
 * stream :2971-3010; clamp_velocity_mach :3312-3335; collide_mrt cumulant_kernel.py:28-209.
 * ---------------------------------------------------------------------------------------- */
static void srt_cell(const float *f_src, float *f_dst, float *rho, float *ux, float *uy,
                     float *uz, const float *Fx, const float *Fy, const float *Fz, float fl,
                     size_t c, size_t N, float omega, float C_darcy)
{
    const float eps = 1e-6f;
    float r = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
    for (int q = 0; q < Q; ++q) {
        float fq = f_src[(size_t)q * N + c];
        r += fq;
        jx += fq * (float)EX[q];
        jy += fq * (float)EY[q];
        jz += fq * (float)EZ[q];
    }
    float inv_r = 1.0f / (r + eps);
    float fx = Fx[c], fy = Fy[c], fz = Fz[c];
    float ux_raw = jx * inv_r + 0.5f * fx * inv_r;
    float uy_raw = jy * inv_r + 0.5f * fy * inv_r;
    float uz_raw = jz * inv_r + 0.5f * fz * inv_r;
    float om = 1.0f - fl;
    float ck = 1.0f + C_darcy * (om * om) / (fl * (fl * fl) + eps);
    float uxf = ux_raw / ck, uyf = uy_raw / ck, uzf = uz_raw / ck;
    float u2 = uxf * uxf + uyf * uyf + uzf * uzf;
    for (int q = 0; q < Q; ++q) {
        float ex = (float)EX[q], ey = (float)EY[q], ez = (float)EZ[q];
        float eu = ex * uxf + ey * uyf + ez * uzf;
        float wq = WQ[q];
        float feq = wq * r * (1.0f + 3.0f * eu + 4.5f * eu * eu - 1.5f * u2);
        float Fe = ex * fx + ey * fy + ez * fz;
        float S = wq * (1.0f - 0.5f * omega) *
                  (3.0f * (Fe) + 9.0f * (eu * Fe) - 3.0f * (uxf * fx + uyf * fy + uzf * fz));
        float fq = f_src[(size_t)q * N + c];
        f_dst[(size_t)q * N + c] = fq - omega * (fq - feq) + S;
    }
    rho[c] = r;
    ux[c] = uxf;
    uy[c] = uyf;
    uz[c] = uzf;
}

void or_collide_srt(const float *f_src, float *f_dst, float *rho, float *ux, float *uy,
                    float *uz, const float *Fx, const float *Fy, const float *Fz,
                    const float *f_l, const uint8_t *flags, float tau, float omega, float dt_lu,
                    float C_darcy, int nx, int ny, int nz)
{
    (void)tau;
    (void)dt_lu;
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (long cc = 0; cc < (long)N; ++cc) {
        size_t c = (size_t)cc;
        uint8_t f = flags[c];
        if (f == SOLID || f == GAS) continue;
        srt_cell(f_src, f_dst, rho, ux, uy, uz, Fx, Fy, Fz, f_l[c], c, N, omega, C_darcy);
    }
}

void or_collide_srt_variable_tau(const float *f_src, float *f_dst, float *rho, float *ux,
                                 float *uy, float *uz, const float *Fx, const float *Fy,
                                 const float *Fz, const float *f_l, const uint8_t *flags,
                                 const float *tau_field, float C_darcy, int nx, int ny, int nz)
{
    const float eps = 1e-6f;
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (long cc = 0; cc < (long)N; ++cc) {
        size_t c = (size_t)cc;
        uint8_t f = flags[c];
        if (f == SOLID || f == GAS) continue;
        float omega = 1.0f / (tau_field[c] + eps);
        srt_cell(f_src, f_dst, rho, ux, uy, uz, Fx, Fy, Fz, f_l[c], c, N, omega, C_darcy);
    }
}

/* collide_mrt's per-cell map (cumulant_kernel.py:28-209) on the populations f of one cell:
 * post-collision populations into fo, density and velocity into ro/uo/vo/wo */
static void mrt_cell_o(const float f[Q], float fx, float fy, float fz, float fl, float omega_s,
                       float omega_b, float C_darcy, float fo[Q], float *ro, float *uo, float *vo,
                       float *wo)
{
    const float cs2 = (float)(1.0 / 3.0);
    const float inv_cs2 = 3.0f;
    const float eps = 1e-6f;
    /* truncated-decimal weights of the reference expansion (cumulant_kernel.py:147-202) */
    const float W0 = 0.3333333333f, W1 = 0.0555555556f, W7 = 0.0277777778f;
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
    float om = 1.0f - fl;
    float ck = 1.0f + C_darcy * (om * om) / (fl * (fl * fl) + eps);
    float u = uxr / ck, v = uyr / ck, w = uzr / ck;
    float Pxx = f[1] + f[2] + f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] + f[14];
    float Pyy = f[3] + f[4] + f[7] + f[8] + f[9] + f[10] + f[15] + f[16] + f[17] + f[18];
    float Pzz = f[5] + f[6] + f[11] + f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
    float Pxy = f[7] * 1.0f * 1.0f + f[8] * -1.0f * -1.0f + f[9] * 1.0f * -1.0f +
                f[10] * -1.0f * 1.0f;
    float Pxz = f[11] * 1.0f * 1.0f + f[12] * -1.0f * -1.0f + f[13] * 1.0f * -1.0f +
                f[14] * -1.0f * 1.0f;
    float Pyz = f[15] * 1.0f * 1.0f + f[16] * -1.0f * -1.0f + f[17] * 1.0f * -1.0f +
                f[18] * -1.0f * 1.0f;
    float kxx = Pxx - r * u * u, kyy = Pyy - r * v * v, kzz = Pzz - r * w * w;
    float kxy = Pxy - r * u * v, kxz = Pxz - r * u * w, kyz = Pyz - r * v * w;
    float km = (kxx + kyy + kzz) / 3.0f;
    float kmp = km - omega_b * (km - r * cs2);
    float kxxp = kmp + (1.0f - omega_s) * (kxx - km);
    float kyyp = kmp + (1.0f - omega_s) * (kyy - km);
    float kzzp = kmp + (1.0f - omega_s) * (kzz - km);
    float kxyp = kxy - omega_s * kxy;
    float kxzp = kxz - omega_s * kxz;
    float kyzp = kyz - omega_s * kyz;
    float u2 = u * u + v * v + w * w;
    float uF = u * fx + v * fy + w * fz;
    float dxx = 1.0f * (kxxp - r * cs2), dyy = 1.0f * (kyyp - r * cs2),
          dzz = 1.0f * (kzzp - r * cs2);
    float Kq[Q];
    Kq[0] = 0.0f;
    Kq[1] = Kq[2] = dxx;
    Kq[3] = Kq[4] = dyy;
    Kq[5] = Kq[6] = dzz;
    Kq[7] = Kq[8] = (dxx + dyy) + (2.0f * kxyp);
    Kq[9] = Kq[10] = (dxx + dyy) + (-2.0f * kxyp);
    Kq[11] = Kq[12] = (dxx + dzz) + (2.0f * kxzp);
    Kq[13] = Kq[14] = (dxx + dzz) + (-2.0f * kxzp);
    Kq[15] = Kq[16] = (dyy + dzz) + (2.0f * kyzp);
    Kq[17] = Kq[18] = (dyy + dzz) + (-2.0f * kyzp);
    for (int q = 0; q < Q; ++q) {
        float wq = q == 0 ? W0 : (q < 7 ? W1 : W7);
        float ex = (float)EX[q], ey = (float)EY[q], ez = (float)EZ[q];
        float eu = ex * u + ey * v + ez * w;
        float eF = ex * fx + ey * fy + ez * fz;
        float feq = wq * (r + r * inv_cs2 * (eu) + r * inv_cs2 * inv_cs2 * 0.5f * (eu) * (eu) -
                          r * 0.5f * inv_cs2 * u2 + 0.5f * inv_cs2 * inv_cs2 * (Kq[q]));
        float S = wq * (1.0f - 0.5f * omega_s) *
                  (inv_cs2 * (eF) + inv_cs2 * inv_cs2 * (eF) * (eu) - inv_cs2 * (uF));
        fo[q] = feq + S;
    }
    *ro = r;
    *uo = u;
    *vo = v;
    *wo = w;
}

void or_collide_mrt(const float *f_src, float *f_dst, float *rho, float *ux, float *uy,
                    float *uz, const float *Fx, const float *Fy, const float *Fz,
                    const float *f_l, const uint8_t *flags, float omega_s, float omega_b,
                    float C_darcy, int nx, int ny, int nz)
{
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (long cc = 0; cc < (long)N; ++cc) {
        size_t c = (size_t)cc;
        if (flags[c] == SOLID || flags[c] == GAS) continue;
        float f[Q], fo[Q], r, u, v, w;
        for (int q = 0; q < Q; ++q) f[q] = f_src[(size_t)q * N + c];
        mrt_cell_o(f, Fx[c], Fy[c], Fz[c], f_l[c], omega_s, omega_b, C_darcy, fo, &r, &u, &v, &w);
        for (int q = 0; q < Q; ++q) f_dst[(size_t)q * N + c] = fo[q];
        rho[c] = r;
        ux[c] = u;
        uy[c] = v;
        uz[c] = w;
    }
}

/* Lid-driven cavity step (no counterpart in the reference, whose stream has no moving wall,
 * kernels.py:2982-3010): every FLUID cell pulls from its neighbours (all three axes periodic),
 * half-way bounce-back off wall cells, with Ladd's moving-wall term 6 w_q (e_q . u_w) (wall
 * density 1) off LID cells, then collide_mrt's map without force or Darcy damping. */
void or_cavity_step(const float *f_src, float *f_dst, float *rho, float *ux, float *uy,
                    float *uz, const uint8_t *flags, float omega_s, float omega_b, float uwx,
                    float uwy, float uwz, int nx, int ny, int nz)
{
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] != FLUID) continue;
                float f[Q], fo[Q], r, u, v, w;
                for (int q = 0; q < Q; ++q) {
                    int ni = (i - EX[q] + nx) % nx, nj = (j - EY[q] + ny) % ny;
                    int nk = (k - EZ[q] + nz) % nz;
                    size_t n = IDX(ni, nj, nk);
                    uint8_t fn = flags[n];
                    if (fn == FLUID) {
                        f[q] = f_src[(size_t)q * N + n];
                    } else {
                        float b = f_src[(size_t)OPP[q] * N + c];
                        if (fn == CAVITY_LID) {
                            float wq = q == 0 ? (float)(1.0 / 3.0)
                                              : (q < 7 ? (float)(1.0 / 18.0) : (float)(1.0 / 36.0));
                            float eu = (float)EX[q] * uwx + (float)EY[q] * uwy + (float)EZ[q] * uwz;
                            b = b + 6.0f * wq * eu;
                        }
                        f[q] = b;
                    }
                }
                mrt_cell_o(f, 0.0f, 0.0f, 0.0f, 1.0f, omega_s, omega_b, 0.0f, fo, &r, &u, &v, &w);
                for (int q = 0; q < Q; ++q) f_dst[(size_t)q * N + c] = fo[q];
                rho[c] = r;
                ux[c] = u;
                uy[c] = v;
                uz[c] = w;
            }
}

void or_stream(const float *f_src, float *f_dst, const uint8_t *flags, int nx, int ny, int nz)
{
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                uint8_t flag = flags[c];
                if (flag == SOLID) continue;
                for (int q = 0; q < Q; ++q) {
                    int ni = i - EX[q], nj = j - EY[q], nk = k - EZ[q];
                    int nip = (ni + nx) % nx;
                    int njp = imax(0, imin(nj, ny - 1));
                    int nkp = imax(0, imin(nk, nz - 1));
                    size_t n = IDX(nip, njp, nkp);
                    if (flag == GAS) f_dst[(size_t)q * N + c] = f_src[(size_t)q * N + n];
                    else if (flags[n] == SOLID)
                        f_dst[(size_t)q * N + c] = f_src[(size_t)OPP[q] * N + c];
                    else f_dst[(size_t)q * N + c] = f_src[(size_t)q * N + n];
                }
            }
}

void or_clamp_velocity_mach(float *ux, float *uy, float *uz, const uint8_t *flags, float u_max,
                            int nx, int ny, int nz)
{
    const float eps = 1e-12f;
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] == SOLID || flags[c] == GAS) continue;
        float vx = ux[c], vy = uy[c], vz = uz[c];
        float mag = sqrtf(vx * vx + vy * vy + vz * vz);
        if (mag > u_max) {
            float s = u_max / (mag + eps);
            ux[c] = vx * s;
            uy[c] = vy * s;
            uz[c] = vz * s;
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * Telemetry — telemetry_pool_reduce kernels.py:3424-3468; extract_fl_yz_slice :3471-3481.
 * out[0..7] = n_liquid, T_peak, cool, u2_max, i_sum, i_min, i_max, T_global (float atomics of
 * the reference become exact max/min and an exact integer sum here).
 * ---------------------------------------------------------------------------------------- */
void or_telemetry_pool_reduce(const float *T, const float *f_l, const float *phi,
                              const float *dT_dt, const float *ux, const float *uy,
                              const float *uz, double *out, int nx, int ny, int nz)
{
    long n = 0;
    float T_peak = 0.0f, cool = 0.0f, u2m = 0.0f, i_min = 1.0e9f, i_max = -1.0e9f, T_g = 0.0f;
    double i_sum = 0.0;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                float Tc = T[c];
                if (Tc > T_g) T_g = Tc;
                if (phi[c] > 0.05f) {
                    float cr = -dT_dt[c];
                    if (cr > cool) cool = cr;
                }
                if (f_l[c] > 0.5f) {
                    n += 1;
                    if (Tc > T_peak) T_peak = Tc;
                    float u2 = ux[c] * ux[c] + uy[c] * uy[c] + uz[c] * uz[c];
                    if (u2 > u2m) u2m = u2;
                    i_sum += (double)i;
                    if ((float)i < i_min) i_min = (float)i;
                    if ((float)i > i_max) i_max = (float)i;
                }
            }
    out[0] = (double)n;
    out[1] = T_peak;
    out[2] = cool;
    out[3] = u2m;
    out[4] = i_sum;
    out[5] = i_min;
    out[6] = i_max;
    out[7] = T_g;
}

/* ------------------------------------------------------------------------------------------
 * Porosity tracers — inject_tracers kernels.py:3182-3204 (ti.random is replaced by a
 * counter-based hash of the slot index and a seed), advect_tracers :3206-3255.
 * pos is (max_tracers, 3) float; active int32.
 * ---------------------------------------------------------------------------------------- */
static inline uint32_t hash_u32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

static inline float unit_from_hash(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

void or_inject_tracers(float *pos, int32_t *active, int32_t *head, int max_tracers, float ax,
                       float ay, float az, float sigma_m, int spawn_count, uint32_t seed)
{
    for (int s = 0; s < spawn_count; ++s) {
        int32_t h = *head;
        *head = h + 1;
        int slot = (int)((uint32_t)h % (uint32_t)max_tracers);
        active[slot] = 1;
        uint32_t base = hash_u32(seed ^ (uint32_t)h * 2654435761u);
        float r1 = unit_from_hash(hash_u32(base + 1u));
        float r2 = unit_from_hash(hash_u32(base + 2u));
        float rdx = (r1 * 2.0f - 1.0f) * sigma_m;
        float rdy = (r2 * 2.0f - 1.0f) * sigma_m;
        pos[3 * slot + 0] = ax + rdx;
        pos[3 * slot + 1] = ay + rdy;
        pos[3 * slot + 2] = az;
    }
}

void or_advect_tracers(float *pos, int32_t *active, const float *ux, const float *uy,
                       const float *uz, const float *f_l, const uint8_t *flags, float dx,
                       float dt, int max_tracers, int nx, int ny, int nz)
{
    (void)dt;
    for (int p = 0; p < max_tracers; ++p) {
        if (active[p] != 1) continue;
        float px = pos[3 * p] / dx, py = pos[3 * p + 1] / dx, pz = pos[3 * p + 2] / dx;
        int i = (int)px, j = (int)py, k = (int)pz;
        if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) {
            active[p] = 0;
            continue;
        }
        size_t c = IDX(i, j, k);
        uint8_t flag = flags[c];
        if (flag == GAS) {
            active[p] = 0;
            continue;
        }
        if (f_l[c] < 0.05f || flag == SOLID) {
            active[p] = 2;
            continue;
        }
        pos[3 * p] += ux[c] * dx;
        pos[3 * p + 1] += uy[c] * dx;
        pos[3 * p + 2] += uz[c] * dx;
    }
}

/* ------------------------------------------------------------------------------------------
 * Temperature-dependent material tables — lookup_table_1d kernels.py:1685-1707,
 * refresh_thermal_properties :1710-1772, advect_diffuse_temperature_variable :1775-1836,
 * update_phase_variable_cp :1839-1863, clamp_enthalpy_floor :1937-1950 (per-cell cp_rho),
 * clamp_enthalpy_ceiling_variable_cp :1987-2010, compute_marangoni_force_variable :2220-2269,
 * remelt_hot_solid :2454-2486, solidify_trailing_pool :2524-2570. The boundary losses of the
 * table path (kernels.py:1866-1918) do not read cp_rho_field and equal the scalar kernel's.
 * ---------------------------------------------------------------------------------------- */
#define OR_MAX_KNOTS 8

static float lookup_table_1d(float T, const float *t_arr, const float *v_arr, int n,
                             float fallback)
{
    float result = fallback;
    if (n > 0) {
        if (T <= t_arr[0]) {
            result = v_arr[0];
        } else if (T >= t_arr[n - 1]) {
            result = v_arr[n - 1];
        } else {
            /* every bracketing interval overwrites the result: the last match wins */
            for (int idx = 0; idx < OR_MAX_KNOTS - 1; ++idx) {
                if (idx < n - 1) {
                    float t0 = t_arr[idx], t1 = t_arr[idx + 1];
                    if (T >= t0 && T <= t1) {
                        float w = (T - t0) / (t1 - t0 + 1e-8f);
                        result = v_arr[idx] + w * (v_arr[idx + 1] - v_arr[idx]);
                    }
                }
            }
        }
    }
    return result;
}

/* tables: 8 arrays of OR_MAX_KNOTS floats (cp_T, cp_V, k_T, k_V, mu_T, mu_V, dg_T, dg_V) back
 * to back; counts: n_cp, n_k, n_mu, n_dg. consts: rho, dt, dx, cp_fb, k_fb, mu_fb, dg_fb,
 * force_scale, cp_rho_ref, alpha_ref, dgamma_ref, tau_ref, marangoni_scale. */
void or_refresh_thermal_properties(const float *T, float *cp_rho_f, float *alpha_f,
                                   float *dgamma_f, float *tau_f, const float *tables,
                                   const int *counts, const float *consts, int use_tables,
                                   const uint8_t *flags, int nx, int ny, int nz)
{
    const float *cp_T = tables, *cp_V = tables + 8, *k_T = tables + 16, *k_V = tables + 24;
    const float *mu_T = tables + 32, *mu_V = tables + 40, *dg_T = tables + 48, *dg_V = tables + 56;
    float rho = consts[0], dt = consts[1], dx = consts[2];
    float cp_fb = consts[3], k_fb = consts[4], mu_fb = consts[5], dg_fb = consts[6];
    float force_scale = consts[7], cp_rho_ref = consts[8], alpha_ref = consts[9];
    float dgamma_ref = consts[10], tau_ref = consts[11], mscale = consts[12];
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] == SOLID || flags[c] == GAS) {
            cp_rho_f[c] = cp_rho_ref;
            alpha_f[c] = alpha_ref;
            dgamma_f[c] = dgamma_ref;
            tau_f[c] = tau_ref;
            continue;
        }
        float T_c = T[c];
        if (use_tables == 1) {
            float cp = lookup_table_1d(T_c, cp_T, cp_V, counts[0], cp_fb);
            float k_val = lookup_table_1d(T_c, k_T, k_V, counts[1], k_fb);
            float mu_val = lookup_table_1d(T_c, mu_T, mu_V, counts[2], mu_fb);
            float dgamma = lookup_table_1d(T_c, dg_T, dg_V, counts[3], dg_fb);
            float cp_r = rho * cp;
            float alpha_phys = k_val / (rho * cp + 1e-9f);
            float nu_phys = mu_val / (rho + 1e-9f);
            float nu_lu = nu_phys * dt / (dx * dx);
            cp_rho_f[c] = cp_r;
            alpha_f[c] = alpha_phys * dt / (dx * dx);
            dgamma_f[c] = dgamma * force_scale * mscale;
            tau_f[c] = 3.0f * nu_lu + 0.5f;
        } else {
            cp_rho_f[c] = cp_rho_ref;
            alpha_f[c] = alpha_ref;
            dgamma_f[c] = dgamma_ref * mscale;
            tau_f[c] = tau_ref;
        }
    }
}

void or_advect_diffuse_temperature_variable_serial(float *H, const float *T, const float *ux,
                                                   const float *uy, const float *uz,
                                                   const uint8_t *flags, const float *alpha_f,
                                                   const float *cp_rho_f, float dt, int nx,
                                                   int ny, int nz)
{
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == GAS) continue;
                float dH = thermal_dH(H, T, ux, uy, uz, flags, i, j, k, alpha_f[c], dt,
                                      cp_rho_f[c], nx, ny, nz);
                H[c] += dH;
            }
}

void or_advect_diffuse_temperature_variable_jacobi(const float *H_in, float *H_out,
                                                   const float *T, const float *ux,
                                                   const float *uy, const float *uz,
                                                   const uint8_t *flags, const float *alpha_f,
                                                   const float *cp_rho_f, float dt, int nx,
                                                   int ny, int nz)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == GAS) {
                    H_out[c] = H_in[c];
                    continue;
                }
                float dH = thermal_dH(H_in, T, ux, uy, uz, flags, i, j, k, alpha_f[c], dt,
                                      cp_rho_f[c], nx, ny, nz);
                H_out[c] = H_in[c] + dH;
            }
}

void or_update_phase_variable_cp(const float *H, float *T, float *f_l, const float *cp_rho_f,
                                 float L_rho, float T_s, float T_l, float H_sol, float H_liq,
                                 int nx, int ny, int nz)
{
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        float h = H[c];
        float cp_rho = cp_rho_f[c];
        if (h <= H_sol) {
            f_l[c] = 0.0f;
            T[c] = h / (cp_rho + 1e-9f);
        } else if (h >= H_liq) {
            f_l[c] = 1.0f;
            T[c] = T_l + (h - H_liq) / (cp_rho + 1e-9f);
        } else {
            float fl = (h - H_sol) / L_rho;
            f_l[c] = fl;
            T[c] = T_s + fl * (T_l - T_s);
        }
    }
}

void or_clamp_enthalpy_floor(float *H, const float *cp_rho_f, const uint8_t *flags,
                             float T_floor, int nx, int ny, int nz)
{
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] == GAS) continue;
        float h_min = cp_rho_f[c] * T_floor;
        if (H[c] < h_min) H[c] = h_min;
    }
}

void or_clamp_enthalpy_ceiling_variable_cp(float *H, const uint8_t *flags, const float *cp_rho_f,
                                           float H_liq, float T_l, float T_cap, int nx, int ny,
                                           int nz)
{
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] == GAS) continue;
        float cp_r = cp_rho_f[c];
        float h_cap_liquid = H_liq + cp_r * (T_cap - T_l);
        float h_cap_solid = cp_r * T_cap;
        float h = H[c];
        if (h > H_liq) {
            if (h > h_cap_liquid) H[c] = h_cap_liquid;
        } else {
            if (h > h_cap_solid) H[c] = h_cap_solid;
        }
    }
}

void or_compute_marangoni_force_variable(const float *T, const float *phi, const float *f_l,
                                         float *Fx, float *Fy, float *Fz, const uint8_t *flags,
                                         const float *dgamma_f, int nx, int ny, int nz)
{
    const float eps = 1e-6f, fl_min = 0.05f;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == SOLID || flags[c] == GAS) continue;
                if (f_l[c] < fl_min) continue;
                int ip = imin(i + 1, nx - 1), im = imax(i - 1, 0);
                int jp = imin(j + 1, ny - 1), jm = imax(j - 1, 0);
                int kp = imin(k + 1, nz - 1), km = imax(k - 1, 0);
                float dphx = 0.5f * (phi[IDX(ip, j, k)] - phi[IDX(im, j, k)]);
                float dphy = 0.5f * (phi[IDX(i, jp, k)] - phi[IDX(i, jm, k)]);
                float dphz = 0.5f * (phi[IDX(i, j, kp)] - phi[IDX(i, j, km)]);
                float gm = sqrtf(dphx * dphx + dphy * dphy + dphz * dphz);
                if (gm < eps) continue;
                float nxn = dphx / gm, nyn = dphy / gm, nzn = dphz / gm;
                float dTx = 0.5f * (T[IDX(ip, j, k)] - T[IDX(im, j, k)]);
                float dTy = 0.5f * (T[IDX(i, jp, k)] - T[IDX(i, jm, k)]);
                float dTz = 0.5f * (T[IDX(i, j, kp)] - T[IDX(i, j, km)]);
                float ndot = nxn * dTx + nyn * dTy + nzn * dTz;
                float sx = dTx - ndot * nxn, sy = dTy - ndot * nyn, sz = dTz - ndot * nzn;
                float scale = dgamma_f[c] * gm;
                Fx[c] += scale * sx;
                Fy[c] += scale * sy;
                Fz[c] += scale * sz;
            }
}

void or_remelt_hot_solid(const float *T, float *H, float *f_l, float *phi, uint8_t *flags,
                         const float *cp_rho_f, float T_s, float T_l, float L_rho, int nx, int ny,
                         int nz)
{
    const float margin = 5.0f;
    size_t N = NCELL;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] != SOLID) continue;
        float T_c = T[c];
        if (T_c <= T_s + margin) continue;
        float cp_r = cp_rho_f[c];
        float H_sol = cp_r * T_s;
        float fl;
        if (T_c >= T_l) {
            fl = 1.0f;
            H[c] = cp_r * T_l + L_rho + cp_r * (T_c - T_l);
        } else {
            fl = (T_c - T_s) / (T_l - T_s + 1e-6f);
            H[c] = H_sol + fl * L_rho;
        }
        f_l[c] = fl;
        phi[c] = 1.0f;
        flags[c] = FLUID;
    }
}

void or_solidify_trailing_pool(float *T, float *H, float *f_l, float *phi, uint8_t *flags,
                               float *ux, float *uy, float *uz, const float *cp_rho_f,
                               float arc_i, float arc_j, float arc_k, float dir_x, float dir_y,
                               float dir_z, float lookback, float T_freeze, float T_s, int nx,
                               int ny, int nz)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == GAS) continue;
                if (phi[c] < 0.85f || f_l[c] < 0.05f) continue;
                float di = (float)i - arc_i, dj = (float)j - arc_j, dk = (float)k - arc_k;
                float proj = di * dir_x + dj * dir_y + dk * dir_z;
                if (proj > -lookback) continue;
                if (T[c] >= T_freeze) continue;
                float cp_r = cp_rho_f[c];
                H[c] = cp_r * (T_s - 1.0f);
                T[c] = T_s - 1.0f;
                f_l[c] = 0.0f;
                phi[c] = 1.0f;
                flags[c] = SOLID;
                ux[c] = 0.0f;
                uy[c] = 0.0f;
                uz[c] = 0.0f;
            }
}

/* ------------------------------------------------------------------------------------------
 * Moving window — shift_simulation_window_x kernels.py:2620-2723: per (j, k) row, an ascending
 * serial loop copies cell i + n into cell i (reading cells not yet written), then the last n
 * columns are reset. fl[] holds the 21 scalar fp32 fields in the reference's argument order:
 * T H f_l phi T_max T_prev dT_dt t800 t1100 tsol rho ux uy uz Fx Fy Fz cp_rho alpha dgamma tau.
 * ---------------------------------------------------------------------------------------- */
void or_shift_simulation_window_x(int n_shift, float *f_a, float *f_b, float **fl,
                                  uint8_t *flags, float T_amb, float cp_rho, float rho0,
                                  int nz_solid, int nx, int ny, int nz)
{
    size_t N = NCELL;
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx - n_shift; ++i) {
                size_t d = IDX(i, j, k), s = IDX(i + n_shift, j, k);
                for (int a = 0; a < 21; ++a) fl[a][d] = fl[a][s];
                flags[d] = flags[s];
                for (int q = 0; q < Q; ++q) {
                    f_a[(size_t)q * N + d] = f_a[(size_t)q * N + s];
                    f_b[(size_t)q * N + d] = f_b[(size_t)q * N + s];
                }
            }
    int i0 = nx - n_shift;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                if (i < i0) continue;
                size_t c = IDX(i, j, k);
                fl[0][c] = T_amb;
                fl[1][c] = cp_rho * T_amb;
                fl[2][c] = 0.0f;
                fl[10][c] = rho0;
                fl[11][c] = 0.0f;
                fl[12][c] = 0.0f;
                fl[13][c] = 0.0f;
                fl[14][c] = 0.0f;
                fl[15][c] = 0.0f;
                fl[16][c] = 0.0f;
                fl[5][c] = T_amb;
                fl[6][c] = 0.0f;
                fl[4][c] = T_amb;
                fl[7][c] = 0.0f;
                fl[8][c] = 0.0f;
                fl[9][c] = 0.0f;
                if (k < nz_solid) {
                    flags[c] = SOLID;
                    fl[3][c] = 1.0f;
                } else {
                    flags[c] = GAS;
                    fl[3][c] = 0.0f;
                }
                for (int q = 0; q < Q; ++q) {
                    float feq = WQ[q] * rho0;
                    f_a[(size_t)q * N + c] = feq;
                    f_b[(size_t)q * N + c] = feq;
                }
            }
}

/* ------------------------------------------------------------------------------------------
 * Full-tier surface forces — apply_vapor_recoil kernels.py:995-1030,
 * apply_vapor_recoil_clausius_clapeyron :1052-1090 (with _wf_pressure_to_Fz_lu :1033-1035),
 * apply_gas_shear_stress :1093-1139. Per-cell and order-free.
 * ---------------------------------------------------------------------------------------- */
void or_apply_vapor_recoil(float *Fz, const float *T, const float *phi, const uint8_t *flags,
                           float arc_i, float arc_j, float sigma, float recoil_pa, float T_ref,
                           float dt, float dx, float rho_ref, int nx, int ny, int nz)
{
    const float eps = 1e-6f;
    float inv2s2 = 1.0f / (2.0f * sigma * sigma + eps);
    float F_peak_lu = (recoil_pa / (rho_ref * dx)) * dt * dt / dx;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == SOLID || flags[c] == GAS) continue;
                float dphi_z = phi[IDX(i, j, imin(k + 1, nz - 1))] - phi[IDX(i, j, imax(k - 1, 0))];
                if (fabsf(dphi_z) < 0.05f) continue;
                float di = (float)i - arc_i, dj = (float)j - arc_j;
                float r2 = di * di + dj * dj;
                float t_ratio = fminf_(T[c] / (T_ref + eps), 3.0f);
                Fz[c] += -F_peak_lu * t_ratio * waam_expf(-r2 * inv2s2);
            }
}

void or_apply_vapor_recoil_cc(float *Fz, const float *T, const float *phi, const uint8_t *flags,
                              float arc_i, float arc_j, float sigma, float P_ref, float T_boil,
                              float L_vap, float R_spec, float C_acc, float dt, float dx,
                              float rho_ref, int nx, int ny, int nz)
{
    const float eps = 1e-6f;
    float inv2s2 = 1.0f / (2.0f * sigma * sigma + eps);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == SOLID || flags[c] == GAS) continue;
                float dphi_z = phi[IDX(i, j, imin(k + 1, nz - 1))] - phi[IDX(i, j, imax(k - 1, 0))];
                if (fabsf(dphi_z) < 0.05f) continue;
                float Tc = T[c];
                if (Tc <= T_boil) continue;
                float expo = (L_vap / (R_spec + eps)) * (1.0f / (T_boil + eps) - 1.0f / (Tc + eps));
                expo = fminf_(expo, 12.0f);
                float P_vap = C_acc * P_ref * waam_expf(expo);
                float F_peak_phys = P_vap / (rho_ref * dx);
                float F_peak_lu = F_peak_phys * dt * dt / dx;
                float di = (float)i - arc_i, dj = (float)j - arc_j;
                float r2 = di * di + dj * dj;
                Fz[c] += -F_peak_lu * waam_expf(-r2 * inv2s2);
            }
}

void or_apply_gas_shear_stress(float *Fx, float *Fy, const float *phi, const uint8_t *flags,
                               float arc_i, float arc_j, float sigma, float tau_peak, float dt,
                               float dx, float rho_ref, int nx, int ny, int nz)
{
    const float eps = 1e-6f;
    float inv2s2 = 1.0f / (2.0f * sigma * sigma + eps);
    float F_scale = (1.0f / (rho_ref * dx)) * dt * dt / dx;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == SOLID || flags[c] == GAS) continue;
                float dpx = 0.5f * (phi[IDX(imin(i + 1, nx - 1), j, k)] - phi[IDX(imax(i - 1, 0), j, k)]);
                float dpy = 0.5f * (phi[IDX(i, imin(j + 1, ny - 1), k)] - phi[IDX(i, imax(j - 1, 0), k)]);
                float dpz = 0.5f * (phi[IDX(i, j, imin(k + 1, nz - 1))] - phi[IDX(i, j, imax(k - 1, 0))]);
                float gmag = sqrtf(dpx * dpx + dpy * dpy + dpz * dpz);
                if (gmag < 0.08f) continue;
                float di = (float)i - arc_i, dj = (float)j - arc_j;
                float r2 = di * di + dj * dj;
                float r = sqrtf(r2 + eps);
                float rx = di / r, ry = dj / r;
                float nxn = dpx / gmag, nyn = dpy / gmag;
                float rdotn = rx * nxn + ry * nyn;
                float tx = rx - rdotn * nxn, ty = ry - rdotn * nyn;
                float tmag = sqrtf(tx * tx + ty * ty + eps);
                tx /= tmag;
                ty /= tmag;
                float tau = tau_peak * waam_expf(-r2 * inv2s2);
                float Fmag = tau * F_scale;
                Fx[c] += Fmag * tx;
                Fy[c] += Fmag * ty;
            }
}

/* ------------------------------------------------------------------------------------------
 * Lorentz force — elec_build_sigma kernels.py:1236-1251, elec_inject_arc_source :1260-1286,
 * elec_normalize_source :1289-1301, elec_jacobi_step :1315-1356, elec_compute_J :1359-1381,
 * elec_l1_diff :1384-1396, elec_bin_axial_current :1399-1418, elec_prefix_bins :1421-1432,
 * elec_B_axisymmetric :1435-1477, apply_lorentz_JxB :1480-1513. The reference's float atomic
 * sums are taken as the correctly rounded fp32 of the exact sum (Neumaier-compensated fp64).
 * ---------------------------------------------------------------------------------------- */
typedef struct {
    double s, c;
} or_sum;

static inline void or_sum_add(or_sum *a, double x)
{
    double t = a->s + x;
    if (fabs(a->s) >= fabs(x)) a->c += (a->s - t) + x;
    else a->c += (x - t) + a->s;
    a->s = t;
}

void or_elec_build_sigma(float *sigma, const float *f_l, const uint8_t *flags, float s_liq,
                         float s_sol, int nx, int ny, int nz)
{
    size_t N = NCELL;
    for (size_t c = 0; c < N; ++c)
        sigma[c] = flags[c] == GAS ? 1e-8f : (f_l[c] > 0.55f ? s_liq : s_sol);
}

void or_elec_inject_arc_source(float *source, float arc_i, float arc_j, float arc_k,
                               float sigma_cells, float current_A, float dx, int nx, int ny, int nz)
{
    const float eps = 1e-6f;
    float inv2s2 = 1.0f / (2.0f * sigma_cells * sigma_cells + eps);
    float cell_vol = dx * dx * dx;
    int k_max = imin(nz, (int)(arc_k + sigma_cells * 4.0f) + 1);
    int k_min = imax(0, (int)(arc_k - sigma_cells));
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                if (k < k_min || k > k_max) continue;
                float di = (float)i - arc_i, dj = (float)j - arc_j, dk = (float)k - arc_k;
                float r2 = di * di + dj * dj + 0.25f * dk * dk;
                float w = waam_expf(-r2 * inv2s2);
                source[IDX(i, j, k)] = current_A * w / (cell_vol + eps);
            }
}

void or_elec_normalize_source(float *source, float target, float dx, int nx, int ny, int nz)
{
    size_t N = NCELL;
    float cell_vol = dx * dx * dx;
    or_sum t = {0.0, 0.0};
    for (size_t c = 0; c < N; ++c) or_sum_add(&t, (double)(source[c] * cell_vol));
    float total = (float)(t.s + t.c);
    float scale = target / fmaxf_(total, 1e-12f);
    for (size_t c = 0; c < N; ++c) source[c] *= scale;
}

static inline float or_sigma_at(const float *s, int i, int j, int k, int nx, int ny, int nz)
{
    return s[IDX(imax(0, imin(nx - 1, i)), imax(0, imin(ny - 1, j)), imax(0, imin(nz - 1, k)))];
}

void or_elec_jacobi_step(const float *phi_in, float *phi_out, const float *sigma,
                         const float *source, const uint8_t *flags, float dx, float omega,
                         int nz_solid, int nx, int ny, int nz)
{
    const float eps = 1e-12f;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == GAS || (k <= nz_solid && flags[c] == SOLID)) {
                    phi_out[c] = 0.0f;
                    continue;
                }
                float sp = or_sigma_at(sigma, i + 1, j, k, nx, ny, nz);
                float sm = or_sigma_at(sigma, i - 1, j, k, nx, ny, nz);
                float sn = or_sigma_at(sigma, i, j + 1, k, nx, ny, nz);
                float ss = or_sigma_at(sigma, i, j - 1, k, nx, ny, nz);
                float st = or_sigma_at(sigma, i, j, k + 1, nx, ny, nz);
                float sb = or_sigma_at(sigma, i, j, k - 1, nx, ny, nz);
                float denom = sp + sm + sn + ss + st + sb + eps;
                float pn = sp * phi_in[IDX(imin(i + 1, nx - 1), j, k)] +
                           sm * phi_in[IDX(imax(i - 1, 0), j, k)] +
                           sn * phi_in[IDX(i, imin(j + 1, ny - 1), k)] +
                           ss * phi_in[IDX(i, imax(j - 1, 0), k)] +
                           st * phi_in[IDX(i, j, imin(k + 1, nz - 1))] +
                           sb * phi_in[IDX(i, j, imax(k - 1, 0))];
                float phi_new = (pn + source[c] * dx * dx) / denom;
                phi_out[c] = (1.0f - omega) * phi_in[c] + omega * phi_new;
            }
}

void or_elec_l1_diff(const float *a, const float *b, float *out, int nx, int ny, int nz)
{
    size_t N = NCELL;
    or_sum d = {0.0, 0.0}, n = {0.0, 0.0};
    for (size_t c = 0; c < N; ++c) {
        or_sum_add(&d, (double)fabsf(a[c] - b[c]));
        or_sum_add(&n, (double)fabsf(b[c]));
    }
    out[0] = (float)(d.s + d.c);
    out[1] = (float)(n.s + n.c);
}

void or_elec_compute_J(float *Jx, float *Jy, float *Jz, const float *phi, const float *sigma,
                       const uint8_t *flags, float dx, int nx, int ny, int nz)
{
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                if (flags[c] == GAS) {
                    Jx[c] = Jy[c] = Jz[c] = 0.0f;
                    continue;
                }
                float sig = sigma[c];
                float gx = (phi[IDX(imin(i + 1, nx - 1), j, k)] - phi[IDX(imax(i - 1, 0), j, k)]) / (2.0f * dx);
                float gy = (phi[IDX(i, imin(j + 1, ny - 1), k)] - phi[IDX(i, imax(j - 1, 0), k)]) / (2.0f * dx);
                float gz = (phi[IDX(i, j, imin(k + 1, nz - 1))] - phi[IDX(i, j, imax(k - 1, 0))]) / (2.0f * dx);
                Jx[c] = -sig * gx;
                Jy[c] = -sig * gy;
                Jz[c] = -sig * gz;
            }
}

void or_elec_bin_axial_current(const float *Jz, float *bins, float arc_i, float arc_j, float dx,
                               float dr, int n_bins, int nx, int ny, int nz)
{
    or_sum *acc = (or_sum *)calloc((size_t)nz * n_bins, sizeof(or_sum));
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            float di = (float)i - arc_i, dj = (float)j - arc_j;
            float r = sqrtf(di * di + dj * dj);
            int b = imin((int)(r / dr), n_bins - 1);
            for (int k = 0; k < nz; ++k)
                or_sum_add(&acc[(size_t)k * n_bins + b], (double)(Jz[IDX(i, j, k)] * dx * dx));
        }
    for (size_t t = 0; t < (size_t)nz * n_bins; ++t) bins[t] = (float)(acc[t].s + acc[t].c);
    free(acc);
}

void or_elec_prefix_bins(float *bins, int n_bins, int nz)
{
    for (int k = 0; k < nz; ++k) {
        float acc = 0.0f;
        for (int b = 0; b < n_bins; ++b) {
            acc += bins[(size_t)k * n_bins + b];
            bins[(size_t)k * n_bins + b] = acc;
        }
    }
}

void or_elec_B_axisymmetric(float *Bx, float *By, float *Bz, const float *bins, float arc_i,
                            float arc_j, float dx, float dr, float mu0, int n_bins, int nx,
                            int ny, int nz)
{
    const float two_pi = 2.0f * 3.14159265358979323846f;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                float di = (float)i - arc_i, dj = (float)j - arc_j;
                float r_cells = sqrtf(di * di + dj * dj);
                Bx[c] = By[c] = Bz[c] = 0.0f;
                if (r_cells > 1e-3f) {
                    int b = imin((int)(r_cells / dr), n_bins - 1);
                    float I_enc = bins[(size_t)k * n_bins + b];
                    float r_m = r_cells * dx;
                    float B_theta = mu0 * I_enc / (two_pi * r_m);
                    Bx[c] = -B_theta * dj / r_cells;
                    By[c] = B_theta * di / r_cells;
                }
            }
}

void or_apply_lorentz_JxB(float *Fx, float *Fy, float *Fz, const float *Jx, const float *Jy,
                          const float *Jz, const float *Bx, const float *By, const float *Bz,
                          const float *f_l, const uint8_t *flags, float dt, float dx, float rho,
                          int nx, int ny, int nz)
{
    size_t N = NCELL;
    for (size_t c = 0; c < N; ++c) {
        if (flags[c] == GAS || f_l[c] < 0.05f) continue;
        float fx = Jy[c] * Bz[c] - Jz[c] * By[c];
        float fy = Jz[c] * Bx[c] - Jx[c] * Bz[c];
        float fz = Jx[c] * By[c] - Jy[c] * Bx[c];
        float flw = fminf_(1.0f, f_l[c]);
        float scale = flw * dt * dt / (rho * dx);
        Fx[c] += fx * scale;
        Fy[c] += fy * scale;
        Fz[c] += fz * scale;
    }
}

/* compute_vorticity_magnitude kernels.py:3380-3419 */
void or_compute_vorticity_magnitude(const float *ux, const float *uy, const float *uz,
                                    const uint8_t *flags, float *out, float dt, int nx, int ny,
                                    int nz)
{
    float scale = 1.0f / (dt + 1e-12f);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k);
                out[c] = 0.0f;
                if (flags[c] == GAS) continue;
                int ip = imin(i + 1, nx - 1), im = imax(i - 1, 0);
                int jp = imin(j + 1, ny - 1), jm = imax(j - 1, 0);
                int kp = imin(k + 1, nz - 1), km = imax(k - 1, 0);
                float duz_dy = 0.5f * (uz[IDX(i, jp, k)] - uz[IDX(i, jm, k)]);
                float duy_dz = 0.5f * (uy[IDX(i, j, kp)] - uy[IDX(i, j, km)]);
                float dux_dz = 0.5f * (ux[IDX(i, j, kp)] - ux[IDX(i, j, km)]);
                float duz_dx = 0.5f * (uz[IDX(ip, j, k)] - uz[IDX(im, j, k)]);
                float duy_dx = 0.5f * (uy[IDX(ip, j, k)] - uy[IDX(im, j, k)]);
                float dux_dy = 0.5f * (ux[IDX(i, jp, k)] - ux[IDX(i, jm, k)]);
                float wx = (duz_dy - duy_dz) * scale;
                float wy = (dux_dz - duz_dx) * scale;
                float wz = (duy_dx - dux_dy) * scale;
                out[c] = sqrtf(wx * wx + wy * wy + wz * wz);
            }
}
