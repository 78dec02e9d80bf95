/*
 * waam_common.h — shared definitions for the gfx950 kernels of the coupled step.
 *
 * Numerics contract (all kernels): fp32 storage and arithmetic, one IEEE operation per
 * source operation (built with -ffp-contract=off, correctly rounded fp32 divide and sqrt,
 * denormals preserved), operands in the reference kernel's evaluation order, exp() from
 * waam_math.h. Under that contract a kernel's output is a pure function of its inputs and
 * matches an fp32 CPU evaluation of the same expressions bit for bit.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <initializer_list>

#include "waam_math.h"
#include "../../include/waam_hip.h"

#pragma clang fp contract(off)

namespace waam {

constexpr uint8_t FLUID = WAAM_FLAG_FLUID;
constexpr uint8_t SOLID = WAAM_FLAG_SOLID;
constexpr uint8_t GAS = WAAM_FLAG_GAS;
constexpr uint8_t IFACE = WAAM_FLAG_IFACE;
constexpr int Q = 19;

/* D3Q19 velocity set (reference grid.py:19-34). */
__host__ __device__ constexpr int ex(int q)
{
    constexpr int t[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    return t[q];
}
__host__ __device__ constexpr int ey(int q)
{
    constexpr int t[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    return t[q];
}
__host__ __device__ constexpr int ez(int q)
{
    constexpr int t[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    return t[q];
}
__host__ __device__ constexpr int opp(int q)
{
    constexpr int t[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
    return t[q];
}
/* Equilibrium weights as fp32 constants: 1/3, 1/18, 1/36 rounded from double. */
__host__ __device__ constexpr float wq(int q)
{
    return q == 0 ? (float)(1.0 / 3.0) : (q < 7 ? (float)(1.0 / 18.0) : (float)(1.0 / 36.0));
}

/* Field layouts (every field of a grid shares one):
 *   kLayoutXYZ  (nx, ny, nz) C order, k fastest: the reference's layout;
 *   kLayoutZYX  (nz, ny, nx) C order, i fastest: rows along the x axis, so a wavefront covers
 *               64 cells of one horizontal row and a flat free surface or a SOLID plate fills
 *               whole wavefronts instead of crossing every one of them. */
constexpr int kLayoutXYZ = WAAM_LAYOUT_XYZ;
constexpr int kLayoutZYX = WAAM_LAYOUT_ZYX;

struct Dims {
    int nx, ny, nz;
    int layout;
    __host__ __device__ size_t n() const { return (size_t)nx * (size_t)ny * (size_t)nz; }
    __host__ __device__ size_t idx(int i, int j, int k) const
    {
        return layout == kLayoutZYX ? ((size_t)k * (size_t)ny + (size_t)j) * (size_t)nx + (size_t)i
                                    : ((size_t)i * (size_t)ny + (size_t)j) * (size_t)nz + (size_t)k;
    }
    /* (i, j, k) of linear index c */
    __host__ __device__ void ijk(size_t c, int &i, int &j, int &k) const
    {
        if (layout == kLayoutZYX) {
            i = (int)(c % (size_t)nx);
            size_t r = c / (size_t)nx;
            j = (int)(r % (size_t)ny);
            k = (int)(r / (size_t)ny);
        } else {
            k = (int)(c % (size_t)nz);
            size_t r = c / (size_t)nz;
            j = (int)(r % (size_t)ny);
            i = (int)(r / (size_t)ny);
        }
    }
    /* extent of the contiguous axis and of the slowest axis */
    __host__ __device__ int fast() const { return layout == kLayoutZYX ? nx : nz; }
    __host__ __device__ int slow() const { return layout == kLayoutZYX ? nz : nx; }
#ifdef __HIPCC__
    /* cell of this thread in a grid3() launch (x: contiguous axis, y: j, z: slowest axis);
     * false for the threads past the end of a row */
    __device__ __forceinline__ bool cell(int &i, int &j, int &k, int z0 = 0) const
    {
        int a = blockIdx.x * blockDim.x + threadIdx.x;
        j = blockIdx.y * blockDim.y + threadIdx.y;
        int b = (int)blockIdx.z - z0;  // z0: grid slices ahead of the cells (other work)
        i = layout == kLayoutZYX ? a : b;
        k = layout == kLayoutZYX ? b : a;
        return a < fast() && j < ny;
    }
#endif
};

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
/* ti.min / ti.max on fp32 values that are never NaN on these paths. */
__device__ __forceinline__ float fmn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fmx(float a, float b) { return a > b ? a : b; }

/* Launch helpers ------------------------------------------------------------------------- */
constexpr int kBlock = 256;

inline unsigned grid_for(size_t n, int block = kBlock)
{
    size_t g = (n + (size_t)block - 1) / (size_t)block;
    return (unsigned)(g == 0 ? 1 : g);
}

/* 3-D launch: x over the contiguous axis (64 lanes = one wavefront), y over j, z over the
 * slowest axis (Dims::cell maps a thread to its cell). */
constexpr int kTileFast = 64;
constexpr int kTileJ = 4;
inline dim3 block3() { return dim3(kTileFast, kTileJ, 1); }
inline dim3 grid3(const Dims &d)
{
    return dim3((unsigned)((d.fast() + kTileFast - 1) / kTileFast),
                (unsigned)((d.ny + kTileJ - 1) / kTileJ), (unsigned)d.slow());
}

/* Error plumbing (abi_core.hip) ------------------------------------------------------------ */
int set_error(int code, const char *fmt, ...);
int check_launch(const char *what);
int check_dims(const char *what, int nx, int ny, int nz, int layout);
int check_ptrs(const char *what, std::initializer_list<const void *> ptrs);
inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

/* Device workspace layout (bytes from the start of the caller's scratch buffer). */
constexpr int kReduceBlocks = 1024;
constexpr size_t kScratchPartials = 0;                                  /* 2*kReduceBlocks doubles */
constexpr size_t kScratchTelemetry = kScratchPartials + 2 * kReduceBlocks * sizeof(double);
constexpr size_t kScratchTelemetryParts = kScratchTelemetry + 64;       /* 8*kReduceBlocks doubles */
constexpr size_t kDepScratchOffset = kScratchTelemetryParts + 8 * kReduceBlocks * sizeof(double);
constexpr size_t kDepMaxCells = 512 * 1024;                             /* deposition marks */
/* fused step: per-segment lengths of its two cell lists, one counter per 256-byte line */
constexpr int kListSegments = 256;
constexpr size_t kCounterStride = 256;
constexpr size_t kScratchListCounters = kDepScratchOffset + kDepMaxCells;
constexpr size_t kScratchBytes = kScratchListCounters + 2 * kListSegments * kCounterStride;

/* Compensated addition (Knuth TwoSum): s + e == a + b exactly. */
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

}  // namespace waam

#define WAAM_RETURN_IF(cond) \
    do {                     \
        int _rc = (cond);    \
        if (_rc) return _rc; \
    } while (0)
