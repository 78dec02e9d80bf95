// cavity.hip — the lid-driven cavity step: D3Q19 pull stream over all cells with half-way
// bounce-back and a moving lid, then the central-moment MRT collision.
//
// The reference's stream has bounce-back only (kernels.py:2982-3010) and, with its buffer
// choreography, never transports momentum between fluid cells; the cavity needs both, so this
// step pulls every FLUID cell from its neighbours every step (all three axes periodic: walls
// are cells). A population that would come from a wall cell is the cell's own opposite
// population (half-way bounce-back); off a LID cell it gains Ladd's moving-wall term
// 6 w_q (e_q . u_w) (wall density 1). The collision is collide_mrt's per-cell map
// (cumulant_kernel.py:28-209) with no force and no Darcy damping.
#include "lbm_cell.h"

namespace waam {

__global__ void __launch_bounds__(256) k_cavity_step(const float *__restrict__ src,
                                                     float *__restrict__ dst, float *rho,
                                                     float *ux, float *uy, float *uz,
                                                     const uint8_t *__restrict__ flags,
                                                     float omega_s, float omega_b, float uwx,
                                                     float uwy, float uwz, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    size_t c = d.idx(i, j, k);
    if (flags[c] != FLUID) return;
    size_t N = d.n();
    float fin[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        int ni = i - ex(q), nj = j - ey(q), nk = k - ez(q);
        ni = ni < 0 ? ni + d.nx : (ni >= d.nx ? ni - d.nx : ni);
        nj = nj < 0 ? nj + d.ny : (nj >= d.ny ? nj - d.ny : nj);
        nk = nk < 0 ? nk + d.nz : (nk >= d.nz ? nk - d.nz : nk);
        size_t n = d.idx(ni, nj, nk);
        uint8_t fn = flags[n];
        float pulled = src[(size_t)q * N + n];
        float b = src[(size_t)opp(q) * N + c];
        float eu = (float)ex(q) * uwx + (float)ey(q) * uwy + (float)ez(q) * uwz;
        float moving = b + 6.0f * wq(q) * eu;
        fin[q] = fn == FLUID ? pulled : (fn == WAAM_CAVITY_LID ? moving : b);
    }
    CellOut o;
    mrt_cell(fin, 0.0f, 0.0f, 0.0f, 1.0f, omega_s, omega_b, 0.0f, o);
#pragma unroll
    for (int q = 0; q < Q; ++q) dst[(size_t)q * N + c] = o.f[q];
    rho[c] = o.rho;
    ux[c] = o.u;
    uy[c] = o.v;
    uz[c] = o.w;
}

}  // namespace waam

using namespace waam;

extern "C" int waam_cavity_step(const float *f_src, float *f_dst, float *rho, float *ux,
                                float *uy, float *uz, const uint8_t *flags, float omega_s,
                                float omega_b, float uwx, float uwy, float uwz, int nx, int ny,
                                int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_cavity_step", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_cavity_step", {f_src, f_dst, rho, ux, uy, uz, flags}));
    if (f_src == f_dst) return set_error(WAAM_ERR_INVALID, "waam_cavity_step: f_src aliases f_dst");
    Dims d{nx, ny, nz, layout};
    hipLaunchKernelGGL(k_cavity_step, grid3(d), block3(), 0, as_stream(stream), f_src, f_dst,
                       rho, ux, uy, uz, flags, omega_s, omega_b, uwx, uwy, uwz, d);
    return check_launch("waam_cavity_step");
}
