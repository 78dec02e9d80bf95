// ctwd.hip — top of the deposited bead for the contact-tip-to-work-distance update
// (reference physics/electrical_stickout.py:39-59: measure_bead_height_m reads every flag back to
// the host each step). Here a device reduction leaves one int for the host to read.
#include "waam_common.h"

namespace waam {

__global__ void k_bead_top_init(int32_t *out) { out[0] = -1; }

/* out[0] = max k over SOLID cells with k >= nz_solid (int max: order-free) */
__global__ void __launch_bounds__(kBlock) k_bead_top(const uint8_t *__restrict__ flags,
                                                     int nz_solid, int32_t *out, Dims d)
{
    int best = -1;
    size_t N = d.n();
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N;
         c += (size_t)gridDim.x * blockDim.x) {
        if (flags[c] != SOLID) continue;
        int i, j, k;
        d.ijk(c, i, j, k);
        if (k >= nz_solid && k > best) best = k;
    }
    for (int off = 32; off > 0; off >>= 1) best = imax(best, __shfl_xor(best, off));
    if ((threadIdx.x & 63) == 0 && best >= 0) atomicMax(out, best);
}

}  // namespace waam

using namespace waam;

extern "C" int waam_bead_top_k(const uint8_t *flags, int nz_solid, int32_t *out, int nx, int ny,
                               int nz, int layout, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_bead_top_k", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_bead_top_k", {flags, out}));
    Dims d{nx, ny, nz, layout};
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_bead_top_init, dim3(1), dim3(1), 0, s, out);
    unsigned g = grid_for(d.n());
    if (g > 4096u) g = 4096u;
    hipLaunchKernelGGL(k_bead_top, dim3(g), dim3(kBlock), 0, s, flags, nz_solid, out, d);
    return check_launch("waam_bead_top_k");
}
