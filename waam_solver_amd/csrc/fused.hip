// fused.hip — the coupled step in seven launches (waam_step_fused).
//
// Launch sequence for one step (reference solvers/coupled_step.py:73-464):
//   1 k_step_begin           torch-column surface height (arc_k); arc weight partial sums
//     (+ k_step_arc_total    their fixed-order total and the energy per unit weight, one block;
//                            a kernel boundary orders the partials device-wide more cheaply than
//                            a fence in every block, which writes back each XCD's L2)
//   2 k_step_thermal         every cell but FLUID/IFACE: arc + enthalpy ceiling applied on the
//                            fly, conduction, boundary losses, floor, ceiling, phase map, T_max,
//                            cooling rate, donor-cell phi advection + re-initialisation; lists
//                            the FLUID/IFACE cells
//   3 k_step_thermal_surface the same chain plus enthalpy advection for the listed cells
//   4 k_step_vof_flags       contact-angle BC, flag classification, phi/f_l consistency, time
//                            above the HAZ thresholds; lists the cells the collision updates
//   5 k_step_collide         listed cells: CSF + Marangoni + arc pressure + gravity + buoyancy,
//                            force clamp, MRT/SRT collision, Mach clamp
//   6 k_advect_tracers       porosity tracers
//   7 k_step_stream          every other cell: the previous step's pull stream, force reset;
//                            every cell: remelt, trailing-pool and cooled-metal solidification
// On steady steps (the row-segment record below) 2, 4 and 7 walk a list of the row segments
// whose results can change (k_step_quiet_list builds it after 1), the pull runs ahead of 7 in
// two launches of its own (k_step_pull_x2), and in the ZYX layout the plate's cold segments are
// marched along z by k_step_thermal_plate instead of being listed.
// Every cell value is produced by the same fp32 expression sequence as the one-kernel-per-
// reference-kernel path, so both paths give identical bits.
//
// Work split. A wavefront covers 64 cells of one (i, j) column, which on a plate crosses
// SOLID, the free surface and GAS; the surface cells carry nearly all the arithmetic (enthalpy
// advection, CSF, collision). The sweeps over every cell (2, 4, 7) do the cheap per-class work
// and append the surface cells to a compacted index list; kernels 3 and 5 run the expensive
// path on full wavefronts from that list (order-free: each cell's result depends only on the
// kernel's inputs). Separate kernels also keep the sweeps' register count, and so their
// occupancy, independent of the surface arithmetic.
//
// Loads before stores. The field pointers arrive in a struct and carry no alias information,
// so every per-cell routine issues its loads before its first store, and selects between
// loaded values rather than branching around loads (a branch lets the compiler sink the loads
// behind it and serialise them).
//
// Deferred stream. The reference streams populations after each collide (f_dst -> f_src) and
// then swaps, so the next collide overwrites the streamed value on every FLUID/IFACE cell. Here
// the stream of step n is folded into kernel 7 of step n+1: a cell that step n+1 collides
// skips the pull (its value would be overwritten), every other cell that was not SOLID at step
// n pulls from its neighbours with step n's flags (kept in sflags). Populations are read once
// and written once per active cell per step instead of twice.
#include <algorithm>
#include <climits>
#include <cmath>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include <hip/hip_ext.h>

#include "arc_cell.h"
#include "droplet_launch.h"
#include "forces_cell.h"
#include "lbm_cell.h"
#include "pop_store.h"
#include "tables_cell.h"
#include "thermal_cell.h"
#include "vof_cell.h"

namespace waam {

/* blocks of the list kernels 3 and 5 (each grid-strides over its list) */
constexpr int kListBlocks = 1024;
/* list blocks sharing a launch with the thermal sweep: about one per CU, so the sweep keeps
 * most of the slots (measured: 1024 or 512 list blocks crowd it, 64 leave the list long) */
constexpr int kListBlocksWithSweep = 256;
/* the pull of fp32 (16-bit per-cell) populations in a launch of its own rather than inside
 * kernel 7 */
#ifndef WAAM_SEPARATE_PULL32
#define WAAM_SEPARATE_PULL32 1
#endif
constexpr bool kSeparatePull32 = WAAM_SEPARATE_PULL32 != 0;
constexpr bool kSeparatePull16 = false;
/* the pull on 8-byte words per lane where the contiguous extent allows it */
#ifndef WAAM_WIDE_PULL
#define WAAM_WIDE_PULL 1
#endif
constexpr bool kWidePull = WAAM_WIDE_PULL != 0;
/* steps one captured graph may hold */
constexpr int kGraphMaxSteps = 4096;

static_assert(kBlock == kListSegments, "the list prefix scan uses one thread per segment");

/* A cell list is split into kListSegments groups with a counter each, so appends spread over
 * many addresses instead of contending on one. Group g takes the cells of the 64-cell blocks
 * b = c / 64 with b mod kListSegments = g, interleaved so that cells that sit together (the
 * surface of a plate fills a few contiguous slices) spread over every group; group g stores
 * its cells at [start(g), start(g) + count_g), a region exactly as long as the number of grid
 * cells in the group, so the groups tile the n-entry list. */
constexpr int kListBlockCells = 64;
struct CellList {
    int32_t *cells;
    int *counts;  // kListSegments counters, kCounterStride bytes apart
    size_t n;     // grid cells
    __device__ __forceinline__ int &count(int g) const
    {
        return *reinterpret_cast<int *>((char *)counts + (size_t)g * kCounterStride);
    }
    __device__ __forceinline__ static int group(size_t c)
    {
        return (int)((c / kListBlockCells) % (size_t)kListSegments);
    }
    /* every round of kListSegments blocks gives each group one block; the partial last round
     * gives the first groups theirs */
    __device__ __forceinline__ size_t start(int g) const
    {
        const size_t round = (size_t)kListBlockCells * kListSegments;
        const size_t full = n / round, rest = n % round;
        return full * kListBlockCells * (size_t)g +
               std::min(rest, (size_t)kListBlockCells * (size_t)g);
    }
    __device__ __forceinline__ size_t cap(int g) const
    {
        const size_t round = (size_t)kListBlockCells * kListSegments;
        const size_t full = n / round, rest = n % round;
        const size_t lo = (size_t)kListBlockCells * (size_t)g;
        const size_t extra = rest > lo ? std::min(rest - lo, (size_t)kListBlockCells) : 0;
        return full * kListBlockCells + extra;
    }
};

/* what moves with the torch from one step to the next: the arc centre, its column, the robot's
 * height floor for arc_k, the (i, j) box outside which the arc heat is exactly zero and the
 * torch's direction of travel */
struct ArcStep {
    float arc_i, arc_j, k_robot;
    int col_i, col_j;
    int box_i0, box_i1, box_j0, box_j1;
    float dir_x, dir_y, dir_z;
};

struct StepConsts {
    waam_step_params p;
    ArcParams arc;
    PhaseConsts pc;
    LossConsts lc;
    ArcPressure ap;
    Recoil rc;
    GasShear gs;
    float cap_l, cap_s, h_min, inv_dt;
    ArcStep at;            // this step's arc (a direct launch)
    const ArcStep *track;  // or, in a captured step, its entry of the graph's device track
    Dims d;
};

/* the step's arc: by value for a direct launch, from device memory for a captured one (the
 * pointer is uniform, so these are scalar loads) */
__device__ __forceinline__ ArcStep arc_step(const StepConsts &s)
{
    if (s.track) {
        const ArcStep *t = s.track;
        ArcStep a;
        a.arc_i = t->arc_i;
        a.arc_j = t->arc_j;
        a.k_robot = t->k_robot;
        a.col_i = t->col_i;
        a.col_j = t->col_j;
        a.box_i0 = t->box_i0;
        a.box_i1 = t->box_i1;
        a.box_j0 = t->box_j0;
        a.box_j1 = t->box_j1;
        a.dir_x = t->dir_x;
        a.dir_y = t->dir_y;
        a.dir_z = t->dir_z;
        return a;
    }
    return s.at;
}

/* the step's arc_k: the column scan of kernel 1 (tot[2]), the host's value (override 1), or the
 * droplet step's scan kept from before the deposition changed the column (override 2, tot[4]) */
__device__ __forceinline__ float step_arc_k(const waam_step_params &p, const float *tot)
{
    return p.use_arc_k_override == 2 ? tot[4] : (p.use_arc_k_override ? p.arc_k_override : tot[2]);
}

/* which = 0: thermal list (FLUID/IFACE by the incoming flags), 1: collision list */
__device__ __forceinline__ CellList cell_list(const waam_step_fields &f, size_t n, int which)
{
    CellList l;
    l.cells = which == 0 ? f.list_thermal : f.list_collide;
    l.counts = reinterpret_cast<int *>((char *)f.scratch + kScratchListCounters +
                                       (size_t)which * kListSegments * kCounterStride);
    l.n = n;
    return l;
}

__device__ __forceinline__ bool is_surface_class(uint8_t fl) { return fl == FLUID || fl == IFACE; }

/* Wavefront-aggregated append of cell c for the lanes with want set: one atomic per wavefront
 * and segment touched; positions within a step are unique, their order is not fixed (consumers
 * are order-free). */
__device__ __forceinline__ void list_append(bool want, size_t c, const CellList &l)
{
    int seg = want ? CellList::group(c) : -1;
    int lane = (int)__lane_id();
    for (;;) {
        unsigned long long m = __ballot(seg >= 0);
        if (m == 0ull) return;
        int leader = __builtin_ctzll(m);
        int g = __shfl(seg, leader);
        unsigned long long mm = __ballot(seg == g);
        int base = 0;
        if (lane == leader) base = atomicAdd(&l.count(g), __popcll(mm));
        base = __shfl(base, leader);
        if (seg == g) {
            const size_t r = (size_t)(base + __popcll(mm & ((1ull << lane) - 1ull)));
            if (r < l.cap(g)) l.cells[l.start(g) + r] = (int32_t)c;
            seg = -1;
        }
    }
}

/* Exclusive prefix of the clamped group lengths into s_pref[0..kListSegments]; every thread
 * of the block calls it. Returns the list length. */
__device__ __forceinline__ int list_prefix(const CellList &l, int *s_pref)
{
    int t = threadIdx.y * blockDim.x + threadIdx.x;
    int v = imin(l.count(t), (int)l.cap(t));
    s_pref[t + 1] = v;
    if (t == 0) s_pref[0] = 0;
    __syncthreads();
    for (int off = 1; off < kListSegments; off <<= 1) {
        int add = t + 1 > off ? s_pref[t + 1 - off] : 0;
        __syncthreads();
        s_pref[t + 1] += add;
        __syncthreads();
    }
    return s_pref[kListSegments];
}

/* cell at flat position t of the list (s_pref[lo] <= t < s_pref[lo + 1]) */
__device__ __forceinline__ size_t list_at(const CellList &l, const int *s_pref, int t)
{
    int lo = 0, hi = kListSegments;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (s_pref[mid] <= t) lo = mid;
        else hi = mid;
    }
    return (size_t)l.cells[l.start(lo) + (size_t)(t - s_pref[lo])];
}

/* (i, j, k) of cell c; the fused step limits grids to INT_MAX cells, so 32-bit division */
__device__ __forceinline__ void cell_ijk(size_t c, const Dims &d, int &i, int &j, int &k)
{
    unsigned cc = (unsigned)c, ny = (unsigned)d.ny;
    unsigned nf = (unsigned)d.fast();
    unsigned r = cc / nf;
    int a = (int)(cc - r * nf);
    int b = (int)(r / ny);
    j = (int)(r - (unsigned)b * ny);
    i = d.layout == kLayoutZYX ? a : b;
    k = d.layout == kLayoutZYX ? b : a;
}

/* Quiet row segments. A segment is the 64 cells one wavefront of a grid3() launch covers along
 * the contiguous axis at one (j, slowest-axis) position. A segment is quiet at the end of a step
 * when every cell in it is GAS with phi exactly 0. For such a cell, kernels 2, 4 and 7 change
 * nothing as long as every cell their stencils read around it is quiet too: the GAS branch of
 * the thermal chain keeps H and maps it to the same T and f_l, the advected and contact-ruled
 * phi stay 0, the flags stay GAS, f_l is reset to 0, no force was set, and no HAZ or freeze rule
 * applies to GAS. A segment that was quiet at the end of the previous step, with all its
 * neighbours, and at the end of the step before as well, therefore holds this step's results in
 * both buffers of every pair already (each buffer was last written by a step that left it
 * quiet): a steady step lists the other, active, segments first (kernel 1b) and runs the three
 * sweeps over that list only. The record: bits[seg] bit 0 = quiet at the end of the last step,
 * bit 1 = quiet at the end of the step before (written by kernel 7 for the segments it visits;
 * a segment left out keeps 3, which is what it would write); then the active-segment count and
 * list.
 *
 * Cold plate. A segment is cold at the end of a step when every cell in it ends the step SOLID,
 * was SOLID after the flag update as well, holds phi 1 and f_l 0, and is below the solidus and
 * both HAZ thresholds. For such a cell kernels 4 and 7 write nothing that differs from what
 * their buffers hold once the segment has been cold for two steps (the flags and stream flags
 * of both buffers are SOLID, phi is 1, no HAZ clock runs, no freeze rule changes a value, the
 * velocity is already zeroed and no force was set): the only per-step change is kernel 2's
 * conduction. bits[seg] bit 2 = cold at the end of the last step, bit 3 = at the end of the
 * step before. Kernel 2 re-checks a listed segment whose bits 2 and 3 are set on its new
 * temperature and liquid fraction; when every cell is still cold it records the segment cold
 * again and leaves it out of kernels 4 and 7 (need[pos] = 0 at its list position). */
struct QuietRec {
    uint8_t *bits;
    int *count;
    int32_t *list;
    uint8_t *need;  // per list position: kernels 4 and 7 visit the segment
    int nbx;        // segments per row
};

__host__ __device__ inline int quiet_nbx(const Dims &d) { return (d.fast() + kTileFast - 1) / kTileFast; }
__host__ __device__ inline size_t quiet_segments(const Dims &d)
{
    return (size_t)quiet_nbx(d) * (size_t)d.ny * (size_t)d.slow();
}
/* byte offsets in the record: bits, the count (own 256-byte line), the list */
__host__ __device__ inline size_t quiet_count_offset(const Dims &d)
{
    return (quiet_segments(d) + 255) / 256 * 256;
}
__host__ __device__ inline size_t quiet_need_offset(const Dims &d)
{
    return quiet_count_offset(d) + 256 + 4 * quiet_segments(d);
}
/* then, 256-aligned, the pull's per-lane masks: one byte per two cells (the second launch of
 * the pull reads its lanes' need and GAS bits there instead of two flag bytes per cell) */
__host__ __device__ inline size_t quiet_pull_offset(const Dims &d)
{
    return (quiet_need_offset(d) + quiet_segments(d) + 255) / 256 * 256;
}
__host__ __device__ inline size_t quiet_bytes(const Dims &d)
{
    return quiet_pull_offset(d) + (d.n() + 1) / 2;
}

/* slices of the plate that kernel 2p marches (ZYX layout with constant material; 0 in the XYZ
 * layout, whose row segments run along z, and with table materials) */
__host__ __device__ inline int plate_slices(const waam_step_params &p, const Dims &d)
{
    int n = p.nz_solid < d.nz ? p.nz_solid : d.nz;
    n = n < 64 ? n : 64;
    return d.layout == kLayoutZYX && !p.use_tables && n > 0 ? n : 0;
}

__device__ __forceinline__ QuietRec quiet_rec(const waam_step_fields &f, const Dims &d)
{
    QuietRec q;
    q.nbx = quiet_nbx(d);
    q.bits = f.quiet;
    q.count = reinterpret_cast<int *>(f.quiet + quiet_count_offset(d));
    q.list = reinterpret_cast<int32_t *>(f.quiet + quiet_count_offset(d) + 256);
    q.need = f.quiet + quiet_need_offset(d);
    return q;
}

__device__ __forceinline__ size_t seg_index(const QuietRec &q, int bx, int j, int b, const Dims &d)
{
    return ((size_t)b * (size_t)d.ny + (size_t)j) * (size_t)q.nbx + (size_t)bx;
}

/* Whether segment (bx, j, b) must be swept this step: not unless bit 0 is set on every segment
 * the stencils of its cells reach and bit 1 on itself. The reach: one cell along the contiguous
 * axis, one row in j, and along the slowest axis two slices below and one above in the ZYX
 * layout (the contact rule reads k-1 and k-2, the phi advection k+1) or one slice either way in
 * XYZ (x; there k-2 lies on the contiguous axis). x wraps, y and z clamp (a clamped neighbour is
 * an edge cell itself). */
__device__ __forceinline__ bool segment_active(const QuietRec &q, int bx, int j, int b,
                                               const Dims &d)
{
    const bool zyx = d.layout == kLayoutZYX;
    /* every byte is requested before any is tested (one memory round trip) */
    unsigned all = 1u;
#pragma unroll
    for (int db = -2; db <= 1; ++db) {
        int bb;
        if (zyx) {
            bb = imax(0, imin(b + db, d.slow() - 1));
        } else {
            bb = db == -2 ? b : b + db;
            bb = bb < 0 ? bb + d.slow() : (bb >= d.slow() ? bb - d.slow() : bb);
        }
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj) {
            const int jj = imax(0, imin(j + dj, d.ny - 1));
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                int x = bx + dx;
                if (zyx) x = x < 0 ? x + q.nbx : (x >= q.nbx ? x - q.nbx : x);
                else x = imax(0, imin(x, q.nbx - 1));
                all &= (unsigned)q.bits[seg_index(q, x, jj, bb, d)];
            }
        }
    }
    const unsigned own = q.bits[seg_index(q, bx, j, b, d)] & 3u;
    return !((own == 3u) && (all & 1u));
}

/* 1b — the active segments of a steady step, one thread per segment, appended to the list a
 * block at a time (one atomic per block; the block's segments in index order, blocks in
 * arrival order: each segment's results depend only on the sweeps' inputs, so the order is
 * free). One atomic per wavefront, on the one counter, took 36 against 20 us at 768x384x192. */
__global__ void __launch_bounds__(kBlock) k_step_quiet_list(waam_step_fields f, StepConsts s)
{
    const Dims &d = s.d;
    const QuietRec q = quiet_rec(f, d);
    const size_t nseg = quiet_segments(d);
    const size_t sg = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int KP = plate_slices(s.p, d);
    bool act = false;
    if (sg < nseg) {
        const int bx = (int)(sg % (size_t)q.nbx);
        const size_t r = sg / (size_t)q.nbx;
        const int b = (int)(r / (size_t)d.ny);
        act = segment_active(q, bx, (int)(r % (size_t)d.ny), b, d);
        /* a plate slice's segment cold at the end of the last two steps goes to kernel 2p,
         * which lists it only if it is no longer cold */
        if (b < KP && (q.bits[sg] & 12u) == 12u) act = false;
    }
    __shared__ int s_cnt[kBlock / 64];
    __shared__ int s_base;
    const unsigned long long m = __ballot(act);
    const int lane = (int)__lane_id();
    const int wave = (int)threadIdx.x >> 6;
    if (lane == 0) s_cnt[wave] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < kBlock / 64; ++w) tot += s_cnt[w];
        s_base = tot ? atomicAdd(q.count, tot) : 0;
    }
    __syncthreads();
    int base = s_base;
    for (int w = 0; w < wave; ++w) base += s_cnt[w];
    if (act) q.list[base + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)sg;
}

/* the cell of this lane in segment seg (block3() blocks: the lanes along the contiguous axis,
 * one wavefront per threadIdx.y) */
__device__ __forceinline__ bool seg_cell(const QuietRec &q, size_t seg, const Dims &d, int &i,
                                         int &j, int &k)
{
    const int bx = (int)(seg % (size_t)q.nbx);
    const size_t r = seg / (size_t)q.nbx;
    j = (int)(r % (size_t)d.ny);
    const int b = (int)(r / (size_t)d.ny);
    const int a = bx * kTileFast + (int)threadIdx.x;
    i = d.layout == kLayoutZYX ? a : b;
    k = d.layout == kLayoutZYX ? b : a;
    return a < d.fast();
}

/* a wavefront's walk over the active-segment list (entries w, w + nw, ... of a grid of nw
 * wavefronts): the entries and their segments' record bytes are loaded 64 at a time, one per
 * lane, and handed out by shuffles, so those loads drop out of each segment's chain of memory
 * round trips. body(segment, list position, record byte). */
template <class Body>
__device__ __forceinline__ void for_listed_segments(const QuietRec &q, int n, Body &&body)
{
    const int nw = (int)(gridDim.x * blockDim.y);
    const int w = (int)(blockIdx.x * blockDim.y + threadIdx.y);
    const int lane = (int)threadIdx.x;
    for (long long t0 = w; t0 < n; t0 += 64LL * nw) {
        const long long tl = t0 + (long long)lane * nw;
        const int mine = tl < n ? q.list[tl] : 0;
        const int mbits = tl < n ? (int)q.bits[mine] : 0;
        const int cnt = (int)std::min<long long>(64, (n - t0 + nw - 1) / nw);
        for (int it = 0; it < cnt; ++it)
            body((size_t)(unsigned)__shfl(mine, it), t0 + (long long)it * nw,
                 (unsigned)__shfl(mbits, it));
    }
}

/* the same walk over the list positions kernel 2 marked needed (the cold plate left out) */
template <class Body>
__device__ __forceinline__ void for_needed_segments(const QuietRec &q, int n, Body &&body)
{
    const int nw = (int)(gridDim.x * blockDim.y);
    const int w = (int)(blockIdx.x * blockDim.y + threadIdx.y);
    const int lane = (int)threadIdx.x;
    for (long long t0 = w; t0 < n; t0 += 64LL * nw) {
        const long long tl = t0 + (long long)lane * nw;
        const bool in = tl < n;
        const int mine = in ? q.list[tl] : 0;
        const bool nd = in && q.need[tl] != 0;
        unsigned long long m = __ballot(nd);
        while (m != 0ull) {
            const int it = __builtin_ctzll(m);
            m &= m - 1ull;
            body((size_t)(unsigned)__shfl(mine, it));
        }
    }
}

/* the wavefront's segment ends the step quiet (`quiet`: this lane's cell is, or it has none)
 * and cold (`cold`, likewise): kernel 7 records both, shifting the previous bits up when the
 * record is continued (a record started afresh holds no cold history) */
__device__ __forceinline__ void record_quiet(const QuietRec &q, size_t seg, bool quiet, bool cold,
                                             bool cont)
{
    const bool all = __ballot(!quiet) == 0ull;
    const bool all_cold = __ballot(!cold) == 0ull;
    if (threadIdx.x == 0) {
        const unsigned old = q.bits[seg];
        q.bits[seg] = (uint8_t)(cont ? (((old & 1u) << 1) | (all ? 1u : 0u) | ((old & 4u) << 1) |
                                        (all_cold ? 4u : 0u))
                                     : (all ? 1u : 0u));
    }
}

/* below the solidus and both HAZ thresholds: no freeze rule and no HAZ clock acts on the cell */
__device__ __forceinline__ float cold_limit(const waam_step_params &p)
{
    return fminf(p.T_s, fminf(p.T800, p.T1100));
}

/* blocks of the sweeps over the active-segment list (each grid-strides over it by wavefront)
 * when the device's residency cannot be queried */
constexpr int kSegBlocks = 2048;

/* the blocks of `threads` threads of `kernel` the device holds at once (its CUs times the
 * blocks per CU its registers allow): a list sweep launched with exactly that many gives every
 * wavefront an equal share of the list from the start, where a larger grid leaves its last
 * blocks to start only as the first ones finish. Queried once per device and kernel. */
template <class K>
static unsigned resident_blocks(K kernel, int threads)
{
    static std::mutex mu;
    static std::map<std::pair<int, const void *>, unsigned> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return kSegBlocks;
    const auto key = std::make_pair(dev, reinterpret_cast<const void *>(kernel));
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    unsigned n = kSegBlocks;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        per_cu > 0 && cus > 0)
        n = (unsigned)(per_cu * cus);
    cache.emplace(key, n);
    return n;
}

/* 1 — surface height in the torch column (reference kernels.py:625-641 and
 *     coupled_step.py:34-59) and per-block compensated arc weight sums (kernels.py:443-496);
 *     empties the collision list, and the thermal list unless it was built ahead. */
__global__ void __launch_bounds__(kBlock) k_step_begin(waam_step_fields f, StepConsts s,
                                                       int do_sums)
{
    __shared__ double sh_s[kBlock], sh_c[kBlock];
    __shared__ int sh_k[kBlock];
    __shared__ float s_arck;
    const Dims &d = s.d;
    const waam_step_params &p = s.p;
    const ArcStep at = arc_step(s);
    if (blockIdx.x == 0) {
        if (!p.thermal_list_ready) cell_list(f, s.d.n(), 0).count(threadIdx.x) = 0;
        cell_list(f, s.d.n(), 1).count(threadIdx.x) = 0;
        if (p.quiet_skip && threadIdx.x == 0) *quiet_rec(f, d).count = 0;
    }
    int best = -1;
    for (int k = threadIdx.x; k < d.nz; k += blockDim.x) {
        size_t c = d.idx(at.col_i, at.col_j, k);
        if (f.flags_in[c] != GAS && f.phi[c] > 0.05f) best = k > best ? k : best;
    }
    sh_k[threadIdx.x] = best;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh_k[threadIdx.x] = imax(sh_k[threadIdx.x], sh_k[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int kf = imax(1, p.nz_solid - 1);
        float ks = (float)(sh_k[0] >= 0 ? sh_k[0] : kf);
        if (ks < 1.0f) ks = (float)imax(1, p.nz_solid - 1);
        if (p.use_k_robot) ks = fmx(ks, at.k_robot);
        s_arck = ks;
        if (blockIdx.x == 0) {
            float *tot = reinterpret_cast<float *>((char *)f.scratch + kScratchTelemetry);
            tot[2] = ks;
        }
    }
    __syncthreads();
    ArcParams ap = s.arc;
    ap.arc_i = at.arc_i;
    ap.arc_j = at.arc_j;
    ap.arc_k = p.use_arc_k_override ? p.arc_k_override : s_arck;
    double sum = 0.0, comp = 0.0;
    size_t N = d.n();
    /* the arc's weights are exactly zero outside its (i, j) box (exp underflows to 0 there), so
     * on a grid wider than the box the sums run over the box's columns only. The TwoSum pairs
     * carry about 100 significand bits, so the total rounded to fp32 does not depend on the
     * order of its terms in practice (it is not order-independent in exact arithmetic: the
     * compensation terms are summed plainly); the flags and liquid fractions of several
     * iterations are loaded together */
    const int bi0 = imax(0, at.box_i0), bi1 = imin(d.nx - 1, at.box_i1);
    const int bj0 = imax(0, at.box_j0), bj1 = imin(d.ny - 1, at.box_j1);
    const unsigned bi = bi1 >= bi0 ? (unsigned)(bi1 - bi0 + 1) : 0u;
    const unsigned bj = bj1 >= bj0 ? (unsigned)(bj1 - bj0 + 1) : 0u;
    const size_t nbox = (size_t)bi * bj * (size_t)d.nz;
    const bool whole = bi == (unsigned)d.nx && bj == (unsigned)d.ny;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    constexpr int kUnroll = 8;
    if (whole) {
        for (size_t base = (size_t)blockIdx.x * blockDim.x + threadIdx.x; do_sums && base < N;
             base += kUnroll * stride) {
            uint8_t flg[kUnroll];
            float flv[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                size_t cell = base + u * stride;
                bool in = cell < N;
                flg[u] = in ? f.flags_in[cell] : (uint8_t)GAS;
                flv[u] = in ? f.fl_in[cell] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                if (flg[u] == GAS) continue;
                int i, j, k;
                cell_ijk(base + u * stride, d, i, j, k);
                float wg, wdep;
                if (!arc_cell_weights(ap, i, j, k, flv[u], wg, wdep)) continue;
                double ss, e;
                two_sum(sum, (double)(wg * wdep), ss, e);
                sum = ss;
                comp += e;
            }
        }
    } else if (d.fast() % 4 == 0) {
        /* the box's rows in groups of four cells along the contiguous axis (the fast extent of
         * the box widened to multiples of four: the cells this adds have weight exactly 0, and
         * adding 0 leaves a TwoSum pair unchanged); one 4-byte flag load per group, and the
         * group's f_l (one 16-byte load) only when a cell of it is not GAS */
        const bool zyx = d.layout == kLayoutZYX;
        const int fa0 = zyx ? (bi0 & ~3) : 0;
        const int fa1 = zyx ? imin(d.nx, (bi1 + 4) & ~3) : d.nz;
        const unsigned ng = (unsigned)imax(0, (fa1 - fa0) >> 2);
        const unsigned nslow = zyx ? (unsigned)d.nz : bi;
        const size_t ngroups = (size_t)ng * bj * nslow;
        constexpr uint32_t kAllGas = 0x01010101u * (uint32_t)GAS;
        constexpr int kG = 4;
        for (size_t base = (size_t)blockIdx.x * blockDim.x + threadIdx.x; do_sums && base < ngroups;
             base += kG * stride) {
            uint32_t fw[kG];
            size_t c0[kG];
#pragma unroll
            for (int u = 0; u < kG; ++u) {
                const size_t t = base + u * stride;
                const bool in = t < ngroups;
                const unsigned tt = in ? (unsigned)t : 0u;
                const unsigned r = tt / ng;
                const int g = (int)(tt - r * ng);
                const unsigned sl = r / bj;
                const int j = bj0 + (int)(r - sl * bj);
                c0[u] = zyx ? d.idx(fa0 + 4 * g, j, (int)sl) : d.idx(bi0 + (int)sl, j, 4 * g);
                fw[u] = in ? *reinterpret_cast<const uint32_t *>(f.flags_in + c0[u]) : kAllGas;
            }
#pragma unroll
            for (int u = 0; u < kG; ++u) {
                if (fw[u] == kAllGas) continue;
                const float4 fl4 = *reinterpret_cast<const float4 *>(f.fl_in + c0[u]);
                const float flv[4] = {fl4.x, fl4.y, fl4.z, fl4.w};
                int i0, j0, k0;
                cell_ijk(c0[u], d, i0, j0, k0);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (((fw[u] >> (8 * e)) & 0xffu) == GAS) continue;
                    const int i = zyx ? i0 + e : i0, k = zyx ? k0 : k0 + e;
                    float wg, wdep;
                    if (!arc_cell_weights(ap, i, j0, k, flv[e], wg, wdep)) continue;
                    double ss, er;
                    two_sum(sum, (double)(wg * wdep), ss, er);
                    sum = ss;
                    comp += er;
                }
            }
        }
    } else {
        /* box cell t: the box's x (ZYX) or z (XYZ) extent fastest; 32-bit arithmetic (the
         * box holds fewer cells than the grid, which the cell lists already bound by 2^31) */
        const bool zyx = d.layout == kLayoutZYX;
        const unsigned ua = zyx ? bi : (unsigned)d.nz;
        for (size_t base = (size_t)blockIdx.x * blockDim.x + threadIdx.x; do_sums && base < nbox;
             base += kUnroll * stride) {
            uint8_t flg[kUnroll];
            float flv[kUnroll];
            unsigned cells[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                size_t t = base + u * stride;
                bool in = t < nbox;
                unsigned tt = in ? (unsigned)t : 0u;
                unsigned r = tt / ua;
                int a = (int)(tt - r * ua);
                unsigned b = r / bj;
                int j = bj0 + (int)(r - b * bj);
                int i = zyx ? bi0 + a : bi0 + (int)b;
                int k = zyx ? (int)b : a;
                cells[u] = (unsigned)d.idx(i, j, k);
                flg[u] = in ? f.flags_in[cells[u]] : (uint8_t)GAS;
                flv[u] = in ? f.fl_in[cells[u]] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                if (flg[u] == GAS) continue;
                int i, j, k;
                cell_ijk((size_t)cells[u], d, i, j, k);
                float wg, wdep;
                if (!arc_cell_weights(ap, i, j, k, flv[u], wg, wdep)) continue;
                double ss, e;
                two_sum(sum, (double)(wg * wdep), ss, e);
                sum = ss;
                comp += e;
            }
        }
    }
    if (!do_sums) return;
    block_reduce_pair(sum, comp, sh_s, sh_c);
    if (threadIdx.x == 0) {
        double *parts = reinterpret_cast<double *>((char *)f.scratch + kScratchPartials);
        parts[2 * blockIdx.x] = sum;
        parts[2 * blockIdx.x + 1] = comp;
    }
}

/* 1b — fixed-order total of the block sums and the energy per unit weight (reference
 *      kernels.py:493-496); one block */
__global__ void __launch_bounds__(kBlock) k_step_arc_total(waam_step_fields f, StepConsts s)
{
    __shared__ double sh_s[kBlock], sh_c[kBlock];
    const double *parts = reinterpret_cast<const double *>((char *)f.scratch + kScratchPartials);
    double sum = 0.0, comp = 0.0;
    for (int b = threadIdx.x; b < kReduceBlocks; b += blockDim.x) {
        double ss, e;
        two_sum(sum, parts[2 * b], ss, e);
        sum = ss;
        comp += parts[2 * b + 1] + e;
    }
    block_reduce_pair(sum, comp, sh_s, sh_c);
    if (threadIdx.x == 0) {
        const waam_step_params &p = s.p;
        float *tot = reinterpret_cast<float *>((char *)f.scratch + kScratchTelemetry);
        float w_total = (float)(sum + comp);
        float epw = 0.0f;
        if (w_total > 1e-9f) epw = p.Q_w * p.eta * p.dt / (w_total * p.dx3);
        tot[0] = w_total;
        tot[1] = epw;
    }
}

/* Enthalpy of a non-GAS cell as the thermal kernel reads it: after this step's arc injection
 * (reference kernels.py:497-510) and enthalpy ceiling (coupled_step.py:108). Holds copies, not
 * pointers into the kernel arguments: the address of an argument forces the compiler to copy
 * the whole argument struct to private memory in some kernels. */
struct HView {
    const float *H_in, *fl_in;
    const uint8_t *flags_in;
    /* temperature-dependent material: the ceiling's cp*rho per cell, from the previous step's
     * refresh (clamp_enthalpy_ceiling_variable_cp, kernels.py:1987-2010) */
    const float *cp_prev;
    Dims d;
    ArcParams ap;
    float epw, H_liq, cap_l, cap_s, T_cap, T_l;
    int box_i0, box_i1, box_j0, box_j1;
    bool arc_on, cap_on, tab;

    /* whether the arc heats cell column (i, j) this step (its weights are exactly 0 outside
     * the box) */
    __device__ __forceinline__ bool heated(int i, int j) const
    {
        return arc_on && i >= box_i0 && i <= box_i1 && j >= box_j0 && j <= box_j1;
    }
    __device__ __forceinline__ float value(float h, float fl, float cp, int i, int j, int k) const
    {
        if (heated(i, j)) {
            float wg, wdep;
            if (arc_cell_weights(ap, i, j, k, fl, wg, wdep)) h += epw * wg * wdep;
        }
        if (cap_on) {
            if (tab) h = ceiling_clamp(h, H_liq, H_liq + cp * (T_cap - T_l), cp * T_cap);
            else h = ceiling_clamp(h, H_liq, cap_l, cap_s);
        }
        return h;
    }
    /* f_l is read only where the arc heats */
    __device__ __forceinline__ float at(size_t c, int i, int j, int k) const
    {
        return value(H_in[c], heated(i, j) ? fl_in[c] : 0.0f, tab ? cp_prev[c] : 0.0f, i, j, k);
    }
    /* _masked_at on the enthalpy view: clamped index, GAS neighbours give the fallback (loads
     * first, select last, so the loads are not sunk behind the flag test) */
    __device__ __forceinline__ float masked(int i, int j, int k, float fallback) const
    {
        int ii = imax(0, imin(i, d.nx - 1));
        int jj = imax(0, imin(j, d.ny - 1));
        int kk = imax(0, imin(k, d.nz - 1));
        size_t c = d.idx(ii, jj, kk);
        uint8_t fg = flags_in[c];
        float v = value(H_in[c], heated(ii, jj) ? fl_in[c] : 0.0f, tab ? cp_prev[c] : 0.0f, ii,
                        jj, kk);
        return fg == GAS ? fallback : v;
    }
};

__device__ __forceinline__ HView make_hview(const waam_step_fields &f, const StepConsts &s)
{
    const float *tot = reinterpret_cast<const float *>((char *)f.scratch + kScratchTelemetry);
    HView hv;
    hv.H_in = f.H_in;
    hv.fl_in = f.fl_in;
    hv.flags_in = f.flags_in;
    hv.d = s.d;
    const ArcStep at = arc_step(s);
    hv.ap = s.arc;
    hv.ap.arc_i = at.arc_i;
    hv.ap.arc_j = at.arc_j;
    hv.ap.arc_k = step_arc_k(s.p, tot);
    hv.epw = tot[1];
    hv.H_liq = s.pc.H_liq;
    hv.cap_l = s.cap_l;
    hv.cap_s = s.cap_s;
    hv.box_i0 = at.box_i0;
    hv.box_i1 = at.box_i1;
    hv.box_j0 = at.box_j0;
    hv.box_j1 = at.box_j1;
    hv.arc_on = s.p.is_welding && !s.p.arc_applied && hv.epw > 0.0f;
    hv.cap_on = s.p.is_welding && !s.p.arc_applied && s.p.enable_cap;
    hv.tab = s.p.use_tables != 0;
    hv.cp_prev = f.cp_rho_prev;
    hv.T_cap = s.p.T_cap;
    hv.T_l = s.p.T_l;
    return hv;
}

/* the step's refresh_thermal_properties values of one cell, stored (temperature-dependent
 * material) */
__device__ __forceinline__ void store_props(const waam_step_fields &f, size_t c,
                                            const CellProps &pr)
{
    f.cp_rho[c] = pr.cp_rho;
    f.alpha_lu[c] = pr.alpha_lu;
    f.dgamma_lu[c] = pr.dgamma_lu;
    f.tau[c] = pr.tau;
}

/* the conduction increment of advect_diffuse_temperature (Jacobi: every neighbour value is
 * pre-kernel, GAS neighbours already replaced by the cell's own T) */
__device__ __forceinline__ float conduction_dH(const CellProps &pr, float T_c, float T_im,
                                               float T_ip, float T_jm, float T_jp, float T_km,
                                               float T_kp)
{
    float lap = T_ip + T_im + T_jp + T_jm + T_kp + T_km - 6.0f * T_c;
    return pr.cp_rho * pr.alpha_lu * lap * 1.0f;
}

/* the rest of a non-GAS cell's chain on its loaded operands: losses through its n_exp exposed
 * faces, floor, ceiling, phase map, HAZ diagnostics (reference kernels.py:3027-3032, 2023-2027;
 * they run before the VOF flag update, so they see this step's incoming flags) and the stores;
 * returns the new T and f_l. fl_zero_held: the cell's f_l output buffer holds 0 already, so a
 * new f_l of 0 is not stored again */
template <bool TAB>
__device__ __forceinline__ float2 thermal_finish(const waam_step_fields &f, const StepConsts &s,
                                                 size_t c, float T_c, float H_c, float dH,
                                                 int n_exp, const CellProps &pr, float T_max,
                                                 float T_prev, bool fl_zero_held)
{
    const waam_step_params &p = s.p;
    float h = H_c + dH;
    if (p.enable_heat_loss) {
        if (n_exp != 0) h -= loss_flux(T_c, s.lc) * (float)n_exp * s.lc.dt / s.lc.dx;
        const float h_min = TAB ? pr.cp_rho * p.T_amb : s.h_min;
        if (h < h_min) h = h_min;
    }
    if (p.enable_cap) {
        if (TAB)
            h = ceiling_clamp(h, s.pc.H_liq, s.pc.H_liq + pr.cp_rho * (p.T_cap - p.T_l),
                              pr.cp_rho * p.T_cap);
        else
            h = ceiling_clamp(h, s.pc.H_liq, s.cap_l, s.cap_s);
    }
    float T, fl;
    PhaseConsts pc = s.pc;
    if (TAB) pc.cp_rho = pr.cp_rho + 1e-9f;
    phase_of(h, pc, T, fl);
    f.H_out[c] = h;
    f.T_out[c] = T;
    if (!(fl_zero_held && fl == 0.0f)) f.fl_out[c] = fl;
    if (T > T_max) f.T_max[c] = T;
    f.dT_dt[c] = (T - T_prev) * s.inv_dt;
    f.T_prev[c] = T;
    if (TAB) store_props(f, c, pr);
    return make_float2(T, fl);
}

/* thermal chain of one cell (reference kernels.py:497-510, 1953-1984, 1611-1682, 2030-2082,
 * 1937-1950, 1519-1556, 3016-3032, 2013-2027 in coupled_step order); returns the cell's new T
 * and f_l.
 * TAB: temperature-dependent material — the cell's properties from the tables at its incoming
 * temperature (the constants on SOLID and GAS cells; reference coupled_step.py:168-182) in
 * place of the constants, the _variable forms of conduction, floor, ceiling and phase map
 * (kernels.py:1775-1863, 1937-1950, 1987-2010), and the properties stored for the rest of the
 * step and the next step's first ceiling.
 * ADV false: the caller passes SOLID and GAS cells only (no enthalpy advection is compiled).
 * fl_zero_held: the cell's f_l output buffer holds 0 already (a SOLID cell of a segment that
 * ended the last two steps cold), so a new f_l of 0 is not stored again */
template <bool TAB, bool ADV = true>
__device__ __forceinline__ float2 thermal_cell(const waam_step_fields &f, const StepConsts &s,
                                             int i, int j, int k, size_t c, uint8_t flag,
                                             float h_gas, bool fl_zero_held = false)
{
    const Dims &d = s.d;
    const waam_step_params &p = s.p;
    if (flag == GAS) {
        /* GAS: only the two phase updates touch it (ceiling, thermal and losses skip GAS);
         * h_gas is the cell's enthalpy, loaded by the caller */
        float h = h_gas, T, fl;
        PhaseConsts pc = s.pc;
        if (TAB) pc.cp_rho = p.refs.cp_rho_ref + 1e-9f;
        phase_of(h, pc, T, fl);
        f.H_out[c] = h;
        f.T_out[c] = T;
        f.fl_out[c] = fl;
        if (TAB) store_props(f, c, cell_props(p.tables, p.refs, 1, GAS, 0.0f));
        return make_float2(T, fl);
    }
    HView hv = make_hview(f, s);
    /* advect_diffuse_temperature (Jacobi: every neighbour value is pre-kernel) */
    float T_c = f.T_in[c];
    float T_im = masked_at(f.T_in, f.flags_in, i - 1, j, k, T_c, d);
    float T_ip = masked_at(f.T_in, f.flags_in, i + 1, j, k, T_c, d);
    float T_jm = masked_at(f.T_in, f.flags_in, i, j - 1, k, T_c, d);
    float T_jp = masked_at(f.T_in, f.flags_in, i, j + 1, k, T_c, d);
    float T_km = masked_at(f.T_in, f.flags_in, i, j, k - 1, T_c, d);
    float T_kp = masked_at(f.T_in, f.flags_in, i, j, k + 1, T_c, d);
    CellProps pr{p.cp_rho, p.alpha_lu, p.dgamma, 0.0f};
    if (TAB) pr = cell_props(p.tables, p.refs, 1, flag, T_c);
    float dH = conduction_dH(pr, T_c, T_im, T_ip, T_jm, T_jp, T_km, T_kp);
    float H_c = hv.at(c, i, j, k);
    if (ADV && flag != SOLID) {
        float u = f.ux[c], v = f.uy[c], w = f.uz[c];
        float H_im = hv.masked(i - 1, j, k, H_c);
        float H_ip = hv.masked(i + 1, j, k, H_c);
        float H_imm = hv.masked(i - 2, j, k, H_im);
        float H_ipp = hv.masked(i + 2, j, k, H_ip);
        float H_jm = hv.masked(i, j - 1, k, H_c);
        float H_jp = hv.masked(i, j + 1, k, H_c);
        float H_jmm = hv.masked(i, j - 2, k, H_jm);
        float H_jpp = hv.masked(i, j + 2, k, H_jp);
        float H_km = hv.masked(i, j, k - 1, H_c);
        float H_kp = hv.masked(i, j, k + 1, H_c);
        float H_kmm = hv.masked(i, j, k - 2, H_km);
        float H_kpp = hv.masked(i, j, k + 2, H_kp);
        float gx = upwind_grad(u, H_imm, H_im, H_c, H_ip, H_ipp);
        float gy = upwind_grad(v, H_jmm, H_jm, H_c, H_jp, H_jpp);
        float gz = upwind_grad(w, H_kmm, H_km, H_c, H_kp, H_kpp);
        dH -= (u * gx + v * gy + w * gz) * 1.0f;
    }
    const int n_exp = p.enable_heat_loss ? exposed_faces(f.flags_in, i, j, k, d) : 0;
    /* HAZ diagnostics' inputs; every load precedes the first store (the field pointers carry
     * no alias information) */
    const float T_max = f.T_max[c], T_prev = f.T_prev[c];
    return thermal_finish<TAB>(f, s, c, T_c, H_c, dH, n_exp, pr, T_max, T_prev, fl_zero_held);
}

/* advect_phi + reinitialize_phi of one cell (reference kernels.py:2320-2391) */
__device__ __forceinline__ float vof_advect_cell(const waam_step_fields &f, const Dims &d, int i,
                                                 int j, int k, uint8_t flag)
{
    float p = advected_phi(f.phi, f.ux, f.uy, f.uz, f.flags_in, i, j, k, d);
    return reinit_phi(p, flag);
}

/* thermal chain and VOF advection of one cell */
template <bool TAB>
__device__ __forceinline__ void thermal_vof_cell(const waam_step_fields &f, const StepConsts &s,
                                                 int i, int j, int k, size_t c, uint8_t flag)
{
    float phi_new = s.p.enable_vof ? vof_advect_cell(f, s.d, i, j, k, flag) : 0.0f;
    thermal_cell<TAB>(f, s, i, j, k, c, flag, f.H_in[c]);
    if (s.p.enable_vof) f.phi_tmp[c] = phi_new;
}

/* 2 — thermal chain and VOF advection of every cell that is not FLUID/IFACE; lists those
 *     (LIST) unless the previous step's kernel 7 already has */
template <bool LIST, bool TAB>
__global__ void __launch_bounds__(256) k_step_thermal(waam_step_fields f, StepConsts s)
{
    const Dims &d = s.d;
    int i, j, k;
    bool in = d.cell(i, j, k);
    size_t c = in ? d.idx(i, j, k) : 0;
    /* the flag, the cell's enthalpy and the phi-advection operands are requested together (a
     * GAS cell, most of a grid, then waits for one memory round trip); a FLUID/IFACE cell's
     * advected phi is computed and dropped (kernel 3 owns those cells) */
    uint8_t flag = in ? f.flags_in[c] : (uint8_t)GAS;
    float h_in = in ? f.H_in[c] : 0.0f;
    float phi_new = (in && s.p.enable_vof) ? vof_advect_cell(f, d, i, j, k, flag) : 0.0f;
    if (LIST) list_append(in && is_surface_class(flag), c, cell_list(f, s.d.n(), 0));
    if (!in || is_surface_class(flag)) return;
    thermal_cell<TAB>(f, s, i, j, k, c, flag, h_in);
    if (s.p.enable_vof) f.phi_tmp[c] = phi_new;
}


/* 2p — on a steady step in the ZYX layout, the listed segments of the plate's slices
 * (k < plate_slices) that ended the last two steps cold: every cell SOLID, so the chain of
 * kernel 2 without flag-dependent work. A block takes 64 cells (i) of kPlateRows adjacent rows
 * (j) and marches up the slices: each lane keeps its cell's T and flag at k-1, k and k+1 in
 * registers, the block stages the rows' T and flags of slice k in LDS for the j neighbours
 * (the rows either side of the block, and the i neighbours beyond the wavefront's 64 cells,
 * are loaded), the i neighbours come from the adjacent lanes, and every load of slice k+1
 * (k+2 for T and flags) is issued before slice k's arithmetic; the slices are split into
 * kPlateChunks ranges, one per block of a column. The arithmetic is kernel 2's
 * (conduction_dH, thermal_finish). Kernel 1b leaves these segments out of the list; this
 * kernel records each cold again (12) or, visited and no longer cold, records 28 and appends it
 * to the list (kernel 2's walk then marks it for kernels 4 and 7, and kernel 7's record drops
 * bit 4). */
constexpr int kPlateRows = 4;
constexpr int kPlateChunks = 3;  // slice ranges per block column (1: 77 us, 3: 72 us)

__global__ void __launch_bounds__(kTileFast * kPlateRows) k_step_thermal_plate(waam_step_fields f,
                                                                              StepConsts s)
{
    __shared__ float sT[kPlateRows][kTileFast];
    __shared__ uint8_t sF[kPlateRows][kTileFast];
    const Dims &d = s.d;
    const waam_step_params &p = s.p;
    const QuietRec q = quiet_rec(f, d);
    const HView hv = make_hview(f, s);
    const float T_cold = cold_limit(p);
    const int KP = plate_slices(p, d);
    const int lane = (int)threadIdx.x, r = (int)threadIdx.y;
    const int bx = (int)blockIdx.x;
    const int j = (int)blockIdx.y * kPlateRows + r;
    const int i = bx * kTileFast + lane;
    const bool row_in = j < d.ny;
    const bool in = row_in && i < d.nx;
    /* every lane loads at clamped indices, so a lane past the grid's edge holds the edge
     * cell's values: exactly what the clamped neighbour of the edge cell reads */
    const int ic = imin(i, d.nx - 1), jc = imin(j, d.ny - 1);
    const bool first = lane == 0, last = lane == kTileFast - 1;
    const int i_h = first ? imax(i - 1, 0) : imin(i + 1, d.nx - 1);  // the lane's halo column
    const bool edge_row = r == 0 || r == kPlateRows - 1;
    const int j_h = r == 0 ? imax(j - 1, 0) : imin(j + 1, d.ny - 1);  // the wave's halo row
    const CellProps pr{p.cp_rho, p.alpha_lu, p.dgamma, 0.0f};
    /* the record bytes of the wave's segments, one slice per lane */
    const unsigned mybits = (row_in && lane < KP) ? q.bits[seg_index(q, bx, jc, lane, d)] : 0u;
    auto hist_at = [&](int k) {
        return ((unsigned)__shfl((int)mybits, k) & 12u) == 12u;
    };
    /* slice state: own T and flag at k-1, k, k+1 (and k+2 in flight), the halo column's and
     * halo row's T and flag at k (and k+1 in flight), the cell's H, f_l, T_max and T_prev at k
     * (and k+1 in flight) */
    auto idx = [&](int ii, int jj, int kk) { return d.idx(ii, jj, imax(0, imin(kk, d.nz - 1))); };
    /* the block's slices: chunk blockIdx.z of kPlateChunks */
    const int kc = (KP + kPlateChunks - 1) / kPlateChunks;
    const int k0 = (int)blockIdx.z * kc, k1 = imin(KP, k0 + kc);
    float Tm, Tc, Tp, Tn;
    uint8_t Fm, Fc, Fp, Fn;
    Tm = f.T_in[idx(ic, jc, k0 - 1)];
    Fm = f.flags_in[idx(ic, jc, k0 - 1)];
    Tc = f.T_in[idx(ic, jc, k0)];
    Fc = f.flags_in[idx(ic, jc, k0)];
    Tp = f.T_in[idx(ic, jc, k0 + 1)];
    Fp = f.flags_in[idx(ic, jc, k0 + 1)];
    float Th = 0.0f, Tr = 0.0f, H = 0.0f, FL = 0.0f, TX = 0.0f, TP = 0.0f;
    uint8_t Fh = GAS, Fr = GAS;
    auto load_k = [&](int k, bool hist, float &th, uint8_t &fh, float &tr, uint8_t &fr,
                      float &h, float &fl, float &tx, float &tp) {
        if (first || last) {
            th = f.T_in[idx(i_h, jc, k)];
            fh = f.flags_in[idx(i_h, jc, k)];
        }
        if (edge_row) {
            tr = f.T_in[idx(ic, j_h, k)];
            fr = f.flags_in[idx(ic, j_h, k)];
        }
        if (hist && in) {
            const size_t c = d.idx(i, j, k);
            h = f.H_in[c];
            fl = hv.heated(i, j) ? f.fl_in[c] : 0.0f;
            tx = f.T_max[c];
            tp = f.T_prev[c];
        }
    };
    if (k0 < k1) load_k(k0, hist_at(k0), Th, Fh, Tr, Fr, H, FL, TX, TP);
    for (int k = k0; k < k1; ++k) {
        const bool hist = hist_at(k);
        /* in flight during slice k's arithmetic: T and flags of k+2, the rest of k+1 */
        Tn = f.T_in[idx(ic, jc, k + 2)];
        Fn = f.flags_in[idx(ic, jc, k + 2)];
        float Th1 = 0.0f, Tr1 = 0.0f, H1 = 0.0f, FL1 = 0.0f, TX1 = 0.0f, TP1 = 0.0f;
        uint8_t Fh1 = GAS, Fr1 = GAS;
        if (k + 1 < k1) load_k(k + 1, hist_at(k + 1), Th1, Fh1, Tr1, Fr1, H1, FL1, TX1, TP1);
        sT[r][lane] = Tc;
        sF[r][lane] = Fc;
        __syncthreads();
        if (hist) {
            /* i neighbours from the adjacent lanes, the wavefront's ends from the halo column
             * (lane i+1 past the grid holds the edge cell: the clamp) */
            float Tim = __shfl(Tc, lane - 1), Tip = __shfl(Tc, lane + 1);
            uint8_t Fim = (uint8_t)__shfl((int)Fc, lane - 1), Fip = (uint8_t)__shfl((int)Fc, lane + 1);
            if (first) {
                Tim = Th;
                Fim = Fh;
            }
            if (last) {
                Tip = Th;
                Fip = Fh;
            }
            /* j neighbours from LDS, the block's edge rows from the halo row (a row past the
             * grid loaded the edge row: the clamp) */
            float Tjm, Tjp;
            uint8_t Fjm, Fjp;
            if (r > 0) {
                Tjm = sT[r - 1][lane];
                Fjm = sF[r - 1][lane];
            } else {
                Tjm = Tr;
                Fjm = Fr;
            }
            if (r < kPlateRows - 1) {
                Tjp = sT[r + 1][lane];
                Fjp = sF[r + 1][lane];
            } else {
                Tjp = Tr;
                Fjp = Fr;
            }
            bool cold = !in;
            if (in) {
                /* masked_at: a GAS neighbour gives the cell's own T */
                const float m_im = Fim == GAS ? Tc : Tim, m_ip = Fip == GAS ? Tc : Tip;
                const float m_jm = Fjm == GAS ? Tc : Tjm, m_jp = Fjp == GAS ? Tc : Tjp;
                const float m_km = Fm == GAS ? Tc : Tm, m_kp = Fp == GAS ? Tc : Tp;
                /* exposed_faces: the in-grid GAS neighbours at k+1, j+-1, i+-1 */
                const int n_exp = p.enable_heat_loss
                                      ? (int)(k + 1 < d.nz && Fp == GAS) +
                                            (int)(j + 1 < d.ny && Fjp == GAS) +
                                            (int)(j - 1 >= 0 && Fjm == GAS) +
                                            (int)(i + 1 < d.nx && Fip == GAS) +
                                            (int)(i - 1 >= 0 && Fim == GAS)
                                      : 0;
                const size_t c = d.idx(i, j, k);
                const float H_c = hv.value(H, FL, 0.0f, i, j, k);
                const float dH = conduction_dH(pr, Tc, m_im, m_ip, m_jm, m_jp, m_km, m_kp);
                const float2 tf = thermal_finish<false>(f, s, c, Tc, H_c, dH, n_exp, pr, TX, TP,
                                                        true);
                cold = tf.x < T_cold && tf.y == 0.0f;
            }
            const bool all_cold = __ballot(!cold) == 0ull;
            if (first && row_in) {
                const size_t seg = seg_index(q, bx, j, k, d);
                q.bits[seg] = all_cold ? 12u : 28u;
                if (!all_cold) q.list[atomicAdd(q.count, 1)] = (int32_t)seg;
            }
        }
        __syncthreads();
        Tm = Tc;
        Fm = Fc;
        Tc = Tp;
        Fc = Fp;
        Tp = Tn;
        Fp = Fn;
        Th = Th1;
        Fh = Fh1;
        Tr = Tr1;
        Fr = Fr1;
        H = H1;
        FL = FL1;
        TX = TX1;
        TP = TP1;
    }
}

/* 2 over the active segments of a steady step (the surface cells go to kernel 3); marks which
 * of them kernels 4 and 7 visit: all but the segments that stay cold plate */
template <bool TAB>
__global__ void __launch_bounds__(256) k_step_thermal_seg(waam_step_fields f, StepConsts s)
{
    const Dims &d = s.d;
    const QuietRec q = quiet_rec(f, d);
    const float T_cold = cold_limit(s.p);
    const int KP = plate_slices(s.p, d);
    const size_t plate_row = (size_t)q.nbx * (size_t)d.ny;  // segments per slice
    for_listed_segments(q, *q.count, [&](size_t seg, long long pos, unsigned bits) {
        int i, j, k;
        const bool in = seg_cell(q, seg, d, i, j, k);
        const size_t c = in ? d.idx(i, j, k) : 0;
        /* cold at the end of the last two steps: every cell is SOLID (so no flag load and no
         * phi advection), its phi_tmp holds 1 (this sweep stored it at the step before) and
         * both f_l buffers hold 0 */
        const bool hist = (bits & 12u) == 12u;
        bool cold = !in;
        /* a plate slice's cold segment: kernel 2p ran it and recorded whether it is still
         * cold (12) or not (28) */
        if (hist && (int)(seg / plate_row) < KP) {
            if (threadIdx.x == 0) q.need[pos] = (bits & 16u) ? 1 : 0;
            return;
        }
        if (hist) {
            if (in) {
                const float2 tf = thermal_cell<TAB, false>(f, s, i, j, k, c, SOLID, 0.0f, true);
                cold = tf.x < T_cold && tf.y == 0.0f;
            }
        } else {
            const uint8_t flag = in ? f.flags_in[c] : (uint8_t)GAS;
            const float h_in = in ? f.H_in[c] : 0.0f;
            /* a SOLID cell's advected phi is re-initialised to 1 whatever its neighbours hold
             * (the plate's segments skip those loads), and kernel 3 owns the surface cells:
             * only GAS cells advect phi here */
            float phi_new = 1.0f;
            if (in && s.p.enable_vof && flag == GAS)
                phi_new = vof_advect_cell(f, d, i, j, k, flag);
            if (in && !is_surface_class(flag)) {
                /* flag is SOLID or GAS: no enthalpy advection */
                const float2 tf = thermal_cell<TAB, false>(f, s, i, j, k, c, flag, h_in);
                if (s.p.enable_vof) f.phi_tmp[c] = phi_new;
                cold = flag == SOLID && tf.x < T_cold && tf.y == 0.0f;
            }
        }
        /* still cold on the new T and f_l */
        const bool skip = hist && __ballot(!cold) == 0ull;
        if (threadIdx.x == 0) {
            q.need[pos] = skip ? 0 : 1;
            if (skip) q.bits[seg] = 12u;
        }
    });
}

/* thermal chain (with enthalpy advection) and VOF advection of the listed FLUID/IFACE cells;
 * worker b of nb, 256 threads each */
template <bool TAB>
__device__ __forceinline__ void thermal_surface_list(const waam_step_fields &f,
                                                     const StepConsts &s, int b, int nb)
{
    __shared__ int s_pref[kListSegments + 1];
    const Dims &d = s.d;
    CellList l = cell_list(f, s.d.n(), 0);
    int n = list_prefix(l, s_pref);
    int tid = threadIdx.y * blockDim.x + threadIdx.x;
    for (int t = b * kBlock + tid; t < n; t += nb * kBlock) {
        size_t c = list_at(l, s_pref, t);
        int i, j, k;
        cell_ijk(c, d, i, j, k);
        thermal_vof_cell<TAB>(f, s, i, j, k, c, f.flags_in[c]);
    }
}

/* 3 — the listed FLUID/IFACE cells (list built by kernel 2) */
template <bool TAB>
__global__ void __launch_bounds__(256) k_step_thermal_surface(waam_step_fields f, StepConsts s)
{
    thermal_surface_list<TAB>(f, s, blockIdx.x, gridDim.x);
}

/* 2+3 in one launch, when the previous step's kernel 7 has listed this step's FLUID/IFACE
 * cells: the first lz grid slices work through the list (dispatched first, so the
 * latency-bound surface cells overlap the bandwidth-bound sweep), the rest sweep every other
 * cell as kernel 2 does. One call site of the cell routine (two inlined copies make the
 * compiler copy the argument structs to private memory). */
template <bool TAB>
__global__ void __launch_bounds__(256) k_step_thermal_all(waam_step_fields f, StepConsts s,
                                                          int lz)
{
    __shared__ int s_pref[kListSegments + 1];
    const Dims &d = s.d;
    const bool list = (int)blockIdx.z < lz;
    CellList l = cell_list(f, s.d.n(), 0);
    int n = 0, t = 0, stride = 0;
    if (list) {
        n = list_prefix(l, s_pref);
        t = (((int)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * kBlock +
            threadIdx.y * blockDim.x + threadIdx.x;
        stride = lz * gridDim.y * gridDim.x * kBlock;
    }
    for (bool first = true;; first = false) {
        int i, j, k;
        size_t c;
        uint8_t flag;
        if (list) {
            if (t >= n) return;
            c = list_at(l, s_pref, t);
            t += stride;
            cell_ijk(c, d, i, j, k);
            flag = f.flags_in[c];
        } else {
            if (!first || !d.cell(i, j, k, lz)) return;
            c = d.idx(i, j, k);
            flag = f.flags_in[c];
        }
        /* as kernel 2: the enthalpy and the phi-advection operands requested with the flag */
        float h_in = f.H_in[c];
        float phi_new = s.p.enable_vof ? vof_advect_cell(f, d, i, j, k, flag) : 0.0f;
        if (!list && is_surface_class(flag)) return;
        thermal_cell<TAB>(f, s, i, j, k, c, flag, h_in);
        if (s.p.enable_vof) f.phi_tmp[c] = phi_new;
    }
}

/* update_time_above_T (reference kernels.py:3261-3283) runs after the VOF flag update and
 * skips cells that are GAS by the updated flags */
struct TimeAbove {
    float T, t800, t1100, tsol;
};
__device__ __forceinline__ TimeAbove time_above_load(const waam_step_fields &f, size_t c)
{
    return TimeAbove{f.T_out[c], f.t800[c], f.t1100[c], f.tsol[c]};
}
__device__ __forceinline__ void time_above_store(const waam_step_fields &f,
                                                 const waam_step_params &p, size_t c,
                                                 const TimeAbove &t)
{
    if (t.T >= p.T800) f.t800[c] = t.t800 + p.dt;
    if (t.T >= p.T1100) f.t1100[c] = t.t1100 + p.dt;
    if (t.T >= p.T_s) f.tsol[c] = t.tsol + p.dt;
}

/* kernel 4's inputs of one cell: its flag, contact-ruled phi, f_l and (for a cell that was not
 * GAS) time-above fields */
struct VofIn {
    uint8_t f_in;
    float p, flo;
    TimeAbove ta;
};

/* the loads of kernel 4 for one cell (no stores: a caller may issue several cells' loads before
 * any cell's stores). SOLID_FAST: a SOLID cell (whose contact-ruled phi is its own) skips the
 * contact rule's neighbour loads, at the price of a second memory round trip for the other
 * cells */
template <bool SOLID_FAST>
__device__ __forceinline__ VofIn vof_flags_load(const waam_step_fields &f, const StepConsts &s,
                                                int i, int j, int k, size_t c)
{
    const Dims &d = s.d;
    /* the flag, the contact rule's operands and f_l are requested together (none depends on
     * another), so a GAS cell waits for one memory round trip; the time-above fields of
     * the cells that were not GAS (nearly all that stay so) follow the flag */
    VofIn v{};
    v.f_in = f.flags_in[c];
    if (s.p.enable_vof) {
        v.p = (s.p.enable_wetting && !(SOLID_FAST && v.f_in == SOLID))
                  ? contact_bc_phi(f.phi_tmp, f.flags_in, i, j, k, s.p.cos_t, s.p.sin_t, d)
                  : f.phi_tmp[c];
        v.flo = f.fl_out[c];
    }
    if (v.f_in != GAS) v.ta = time_above_load(f, c);
    return v;
}

/* the rest of kernel 4 for one cell on its loaded inputs; returns its new flag */
__device__ __forceinline__ uint8_t vof_flags_apply(const waam_step_fields &f, const StepConsts &s,
                                                   int k, size_t c, const VofIn &v)
{
    uint8_t nf = GAS;
    if (!s.p.enable_vof) {
        nf = v.f_in;
    } else {
        nf = classify_cell(k, s.p.nz_solid, v.p, v.flo);
        f.phi[c] = v.p;
        if (nf == GAS || v.p < 0.05f) f.fl_out[c] = 0.0f;
        else if (nf == SOLID) f.fl_out[c] = 0.0f;
    }
    f.flags_out[c] = nf;
    if (nf != GAS) {
        const TimeAbove ta = v.f_in != GAS ? v.ta : time_above_load(f, c);
        time_above_store(f, s.p, c, ta);
    }
    return nf;
}

/* the body of kernel 4 for one cell; returns its new flag */
template <bool SOLID_FAST>
__device__ __forceinline__ uint8_t vof_flags_cell(const waam_step_fields &f, const StepConsts &s,
                                                  int i, int j, int k, size_t c)
{
    return vof_flags_apply(f, s, k, c, vof_flags_load<SOLID_FAST>(f, s, i, j, k, c));
}

/* 4 — apply_contact_angle_phi_bc (Jacobi), update_flags_from_phi, sync_phi_liquid_fraction,
 *     update_time_above_T (reference kernels.py:933-991, 2394-2419, 3484-3504, 3261-3283);
 *     lists the cells kernel 5 collides */
__global__ void __launch_bounds__(256) k_step_vof_flags(waam_step_fields f, StepConsts s)
{
    const Dims &d = s.d;
    if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0)  // kernel 2 has consumed the
        cell_list(f, s.d.n(), 0).count(threadIdx.y * blockDim.x + threadIdx.x) = 0;  // list
    int i, j, k;
    bool in = d.cell(i, j, k);
    size_t c = in ? d.idx(i, j, k) : 0;
    uint8_t nf = in ? vof_flags_cell<false>(f, s, i, j, k, c) : (uint8_t)GAS;
    list_append(in && is_surface_class(nf), c, cell_list(f, s.d.n(), 1));
}

/* 4 over the segments kernel 2 left in */
__global__ void __launch_bounds__(256) k_step_vof_flags_seg(waam_step_fields f, StepConsts s)
{
    const Dims &d = s.d;
    if (blockIdx.x == 0)  // kernel 2 has consumed the list
        cell_list(f, d.n(), 0).count(threadIdx.y * blockDim.x + threadIdx.x) = 0;
    const QuietRec q = quiet_rec(f, d);
    for_needed_segments(q, *q.count, [&](size_t seg) {
        int i, j, k;
        const bool in = seg_cell(q, seg, d, i, j, k);
        const size_t c = in ? d.idx(i, j, k) : 0;
        const uint8_t nf = in ? vof_flags_cell<true>(f, s, i, j, k, c) : (uint8_t)GAS;
        list_append(in && is_surface_class(nf), c, cell_list(f, d.n(), 1));
    });
}

/* Pull of the reference stream (kernels.py:2991-3010) with the flags of the step it belongs
 * to (sfl), done by a whole wavefront of a grid3() launch: every lane loads the population
 * at its own fast-axis position in the neighbouring row (one aligned, coalesced load per
 * direction) and takes the value of a direction with a fast-axis component from the adjacent
 * lane; the lanes at the ends of the wavefront and at the domain's wrap (x) or clamp (z) load
 * directly. Every active lane of the wavefront must call it (it shuffles); `need` says whether
 * this lane stores. Values are exactly those of a cell-by-cell pull. Neighbour indices are
 * recomputed for the bounce-back test rather than kept (19 64-bit indices held across the
 * loads cost 38 registers and a wavefront per SIMD). */
template <class E>
__device__ __forceinline__ void pull_rows(const E *__restrict__ src, E *__restrict__ dst,
                                          const uint8_t *__restrict__ sfl, int i, int j, int k,
                                          bool need, uint8_t fc, const Dims &d)
{
    size_t N = d.n();
    size_t c = d.idx(i, j, k);
    const bool zyx = d.layout == kLayoutZYX;
    const int lane = (int)threadIdx.x;  // block3(): one wavefront per threadIdx.y row
    const int a = zyx ? i : k;          // this lane's fast-axis coordinate
    const int nf = d.fast();
    E v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        int ni = i - ex(q), nj = j - ey(q), nk = k - ez(q);
        ni = ni < 0 ? ni + d.nx : (ni >= d.nx ? ni - d.nx : ni);
        nj = imax(0, imin(nj, d.ny - 1));
        nk = imax(0, imin(nk, d.nz - 1));
        const int sh = zyx ? ex(q) : ez(q);  // fast-axis offset of the source
        E al = src[(size_t)q * N + (zyx ? d.idx(i, nj, nk) : d.idx(ni, nj, k))];
        E nb = sh == 0 ? al : shfl_elem(al, lane - sh);
        bool direct = sh != 0 && (lane - sh < 0 || lane - sh > 63 || a - sh < 0 || a - sh >= nf);
        v[q] = direct ? src[(size_t)q * N + d.idx(ni, nj, nk)] : nb;
    }
    if (!need) return;
    if (fc != GAS) {
        /* bounce-back off SOLID neighbours (a GAS destination pulls unconditionally) */
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            int ni = i - ex(q), nj = j - ey(q), nk = k - ez(q);
            ni = ni < 0 ? ni + d.nx : (ni >= d.nx ? ni - d.nx : ni);
            nj = imax(0, imin(nj, d.ny - 1));
            nk = imax(0, imin(nk, d.nz - 1));
            if (sfl[d.idx(ni, nj, nk)] == SOLID) v[q] = src[(size_t)opp(q) * N + c];
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) dst[(size_t)q * N + c] = v[q];
}

/* body forces, collision and Mach clamp of one FLUID/IFACE cell; populations in storage form
 * PD (fp32 arithmetic either way) */
template <int PD>
__device__ __forceinline__ void collide_cell(const waam_step_fields &f, const StepConsts &s,
                                             int i, int j, int k, size_t c)
{
    using PS = PopStore<PD>;
    const typename PS::E *f_src = static_cast<const typename PS::E *>(f.f_src);
    typename PS::E *f_dst = static_cast<typename PS::E *>(f.f_dst);
    const Dims &d = s.d;
    const waam_step_params &p = s.p;
    size_t N = d.n();
    /* every load first (populations, forces, stencils), then the arithmetic */
    float fin[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) fin[q] = PS::load(f_src, (size_t)q * N + c, q);
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if (p.forces_from_fields) {
        fx = f.Fx[c];
        fy = f.Fy[c];
        fz = f.Fz[c];
    }
    float fl = f.fl_out[c], T_c = f.T_out[c];
    ForceStencil fs;
    load_force_stencil(f.phi, f.T_out, f.flags_out, i, j, k, d, fs);
    const float *tot = reinterpret_cast<const float *>((char *)f.scratch + kScratchTelemetry);
    float arc_k = step_arc_k(p, tot);
    /* forces, reference order coupled_step.py:298-362 */
    float tx, ty, tz;
    if (p.enable_csf &&
        csf_force(fs, i, j, k, d, p.gamma_lu, p.enable_wetting, p.cos_t, p.sin_t, tx, ty, tz)) {
        fx += tx;
        fy += ty;
        fz += tz;
    }
    /* temperature-dependent material: the cell's coefficient of this step's refresh
     * (compute_marangoni_force_variable, kernels.py:2220-2269) */
    const float dgamma = p.use_tables ? f.dgamma_lu[c] : p.dgamma;
    if (marangoni_force(fs, fl, dgamma, tx, ty, tz)) {
        fx += tx;
        fy += ty;
        fz += tz;
    }
    const ArcStep at = arc_step(s);
    if (p.is_welding && p.enable_gas_shear) {
        GasShear gs = s.gs;
        gs.arc_i = at.arc_i;
        gs.arc_j = at.arc_j;
        if (gas_shear_force(fs.phi, i, j, gs, tx, ty)) {
            fx += tx;
            fy += ty;
        }
    }
    if (p.is_welding) {
        ArcPressure ap = s.ap;
        ap.arc_i = at.arc_i;
        ap.arc_j = at.arc_j;
        ap.arc_k = arc_k;
        if (arc_pressure_force(fs.phi.c[4], fs.phi.c[5], i, j, k, ap, tz)) fz += tz;
        if (p.enable_recoil) {
            Recoil rc = s.rc;
            rc.arc_i = at.arc_i;
            rc.arc_j = at.arc_j;
            if (recoil_force(fs.phi.c[4], fs.phi.c[5], T_c, i, j, rc, tz)) fz += tz;
        }
    }
    if (p.enable_hydro && fl > 0.0f) fz += -1.0f * p.g_lu * fl;
    if (fl > 0.0f) fz += 1.0f * p.g_lu * p.beta * (T_c - p.T_l) * fl;
    if (p.F_cap > 0.0f) {
        float mag = sqrtf(fx * fx + fy * fy + fz * fz);
        if (mag > p.F_cap) {
            float sc = p.F_cap / (mag + 1e-12f);
            fx = fx * sc;
            fy = fy * sc;
            fz = fz * sc;
        }
    }
    CellOut o;
    if (p.collision == 0) {
        mrt_cell(fin, fx, fy, fz, fl, p.omega_s, p.omega_b, p.C_darcy, o);
    } else {
        /* collision 2: collide_srt_variable_tau (reference coupled_step.py:364-373), the
         * relaxation time of the cell's refresh */
        const float om = p.collision == 2 ? 1.0f / (f.tau[c] + 1e-6f) : p.omega_s;
        srt_cell(fin, fx, fy, fz, fl, om, p.C_darcy, o);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) PS::store(f_dst, (size_t)q * N + c, q, o.f[q]);
    float u = o.u, v = o.v, w = o.w;
    if (p.u_cap > 0.0f) {
        float mag = sqrtf(u * u + v * v + w * w);
        if (mag > p.u_cap) {
            float sc = p.u_cap / (mag + 1e-12f);
            u = u * sc;
            v = v * sc;
            w = w * sc;
        }
    }
    f.rho[c] = o.rho;
    f.ux[c] = u;
    f.uy[c] = v;
    f.uz[c] = w;
    if (p.store_forces) {
        f.Fx[c] = fx;
        f.Fy[c] = fy;
        f.Fz[c] = fz;
    }
}

/* 5 — forces + collision of the FLUID/IFACE cells listed by kernel 4 */
template <int PD>
__global__ void __launch_bounds__(256) k_step_collide(waam_step_fields f, StepConsts s)
{
    __shared__ int s_pref[kListSegments + 1];
    const Dims &d = s.d;
    CellList l = cell_list(f, s.d.n(), 1);
    int n = list_prefix(l, s_pref);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        size_t c = list_at(l, s_pref, t);
        int i, j, k;
        cell_ijk(c, d, i, j, k);
        collide_cell<PD>(f, s, i, j, k, c);
    }
}

/* remelt_hot_solid_scalar, solidify_trailing_pool_scalar, solidify_cooled_metal (reference
 * kernels.py:2489-2520, 2573-2617, 2422-2450) in that order for one non-GAS cell, on loaded
 * values; a field is stored exactly when one of the three writes it, with its last value */
__device__ __forceinline__ uint8_t freeze_cell(const waam_step_fields &f, const StepConsts &s,
                                               int i, int j, int k, size_t c, uint8_t flag,
                                               float T, float phi, float fl, float arc_k)
{
    const waam_step_params &p = s.p;
    bool wH = false, wT = false, wfl = false, wphi = false, wflag = false, wu = false;
    float H = 0.0f;
    /* temperature-dependent material: the cell's cp*rho of this step's refresh (the _variable
     * remelt and trailing rules, kernels.py:2454-2486, 2524-2570) */
    const float cp_rho = p.use_tables ? f.cp_rho[c] : p.cp_rho;
    if (flag == SOLID && T > p.T_s + 5.0f) {
        float H_sol = cp_rho * p.T_s;
        if (T >= p.T_l) {
            fl = 1.0f;
            H = cp_rho * p.T_l + p.L_rho + cp_rho * (T - p.T_l);
        } else {
            fl = (T - p.T_s) / (p.T_l - p.T_s + 1e-6f);
            H = H_sol + fl * p.L_rho;
        }
        phi = 1.0f;
        flag = FLUID;
        wH = wfl = wphi = wflag = true;
    }
    if (p.enable_trailing && !(phi < 0.85f || fl < 0.05f)) {
        const ArcStep at = arc_step(s);
        float di = (float)i - at.arc_i, dj = (float)j - at.arc_j, dk = (float)k - arc_k;
        float proj = di * at.dir_x + dj * at.dir_y + dk * at.dir_z;
        if (!(proj > -p.lookback) && !(T >= p.T_freeze)) {
            H = cp_rho * (p.T_s - 1.0f);
            T = p.T_s - 1.0f;
            fl = 0.0f;
            phi = 1.0f;
            flag = SOLID;
            wH = wT = wfl = wphi = wflag = wu = true;
        }
    }
    if (!(phi < 0.85f) && fl < 0.02f && T < p.T_s) {
        flag = SOLID;
        phi = 1.0f;
        fl = 0.0f;
        wflag = wphi = wfl = true;
        if (p.zero_velocity == 1) wu = true;
    }
    if (wH) f.H_out[c] = H;
    if (wT) f.T_out[c] = T;
    if (wfl) f.fl_out[c] = fl;
    if (wphi) f.phi[c] = phi;
    if (wflag) f.flags_out[c] = flag;
    if (wu) {
        f.ux[c] = 0.0f;
        f.uy[c] = 0.0f;
        f.uz[c] = 0.0f;
    }
    return flag;
}

/* 7 — every cell: stream flags; cells the collision skipped: force reset and the previous
 *     step's pull; then the freeze/remelt updates of every non-GAS cell (cell-local; the pull
 *     reads only populations and stream flags, which they do not touch); lists the cells that
 *     end the step FLUID/IFACE, the next step's thermal list; records which row segments end
 *     the step quiet. Returns whether the cell ends the step quiet. */
template <class E, bool PULL>
__device__ __forceinline__ bool stream_cell(const waam_step_fields &f, const StepConsts &s, int i,
                                            int j, int k)
{
    const Dims &d = s.d;
    const waam_step_params &p = s.p;
    size_t c = d.idx(i, j, k);
    uint8_t fc = f.flags_out[c];
    uint8_t fp = f.sflags_prev[c];
    bool surface = is_surface_class(fc);
    bool freeze = p.enable_freeze && fc != GAS;
    float T = 0.0f, fl = 0.0f;
    const float phi = f.phi[c];
    if (freeze) {
        T = f.T_out[c];
        fl = f.fl_out[c];
    }
    bool need = !surface && p.pending_stream && fp != SOLID;
    if (PULL && __ballot(need) != 0ull)
        pull_rows(static_cast<const E *>(f.f_src), static_cast<E *>(f.f_dst), f.sflags_prev, i, j,
                  k, need, fp, d);
    f.sflags_next[c] = fc;
    /* forces are non-zero only on the cells the previous step collided, unless the step before
     * took them from the fields (then force_reset_all) */
    if (!surface && p.store_forces && !p.forces_from_fields &&
        (p.force_reset_all || is_surface_class(fp))) {
        f.Fx[c] = 0.0f;
        f.Fy[c] = 0.0f;
        f.Fz[c] = 0.0f;
    }
    uint8_t fin = fc;
    if (freeze) {
        const float *tot = reinterpret_cast<const float *>((char *)f.scratch + kScratchTelemetry);
        fin = freeze_cell(f, s, i, j, k, c, fc, T, phi, fl, step_arc_k(p, tot));
    }
    list_append(is_surface_class(fin), c, cell_list(f, s.d.n(), 0));
    return fc == GAS && phi == 0.0f;  // the freeze rules never leave a cell GAS
}

/* kernel 7 without the pull: a lane takes kStreamSlices cells along the slowest axis and issues
 * the flag loads of all of them, then the freeze inputs of all of them, before any of the
 * per-cell work: the sweep moves a few bytes per cell and is otherwise bound by the two
 * dependent memory round trips of each cell */
constexpr int kStreamSlices = 4;

__device__ __forceinline__ bool stream_rest_cell(const waam_step_fields &f, const StepConsts &s,
                                                 int i, int j, int k, size_t c, uint8_t fc,
                                                 uint8_t fp, float T, float phi, float fl)
{
    const waam_step_params &p = s.p;
    bool surface = is_surface_class(fc);
    f.sflags_next[c] = fc;
    if (!surface && p.store_forces && !p.forces_from_fields &&
        (p.force_reset_all || is_surface_class(fp))) {
        f.Fx[c] = 0.0f;
        f.Fy[c] = 0.0f;
        f.Fz[c] = 0.0f;
    }
    uint8_t fin = fc;
    if (p.enable_freeze && fc != GAS) {
        const float *tot = reinterpret_cast<const float *>((char *)f.scratch + kScratchTelemetry);
        fin = freeze_cell(f, s, i, j, k, c, fc, T, phi, fl, step_arc_k(p, tot));
    }
    list_append(is_surface_class(fin), c, cell_list(f, s.d.n(), 0));
    return fc == GAS && phi == 0.0f;  // the freeze rules never leave a cell GAS
}

template <class E, bool PULL>
__global__ void __launch_bounds__(256) k_step_stream(waam_step_fields f, StepConsts s)
{
    const Dims &d = s.d;
    const QuietRec q = quiet_rec(f, d);
    int i, j, k;
    if (PULL) {
        const bool in = d.cell(i, j, k);
        if (j >= d.ny) return;  // the wavefront's whole row is past the grid
        bool quiet = true;
        if (in) quiet = stream_cell<E, true>(f, s, i, j, k);
        record_quiet(q, seg_index(q, (int)blockIdx.x, j, (int)blockIdx.z, d), quiet, false,
                     false);
        return;
    }
    const int a = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    j = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (j >= d.ny) return;  // the wavefront's whole row is past the grid
    const bool in = a < d.fast();
    const bool zyx = d.layout == kLayoutZYX;
    const int b0 = (int)blockIdx.z * kStreamSlices;
    const bool freeze_on = s.p.enable_freeze != 0;
    size_t c[kStreamSlices];
    uint8_t fc[kStreamSlices], fp[kStreamSlices];
    float T[kStreamSlices], phi[kStreamSlices], fl[kStreamSlices];
#pragma unroll
    for (int t = 0; t < kStreamSlices; ++t) {
        const int b = imin(b0 + t, d.slow() - 1);
        c[t] = zyx ? d.idx(in ? a : 0, j, b) : d.idx(b, j, in ? a : 0);
        fc[t] = in ? f.flags_out[c[t]] : (uint8_t)GAS;
        fp[t] = in ? f.sflags_prev[c[t]] : (uint8_t)GAS;
        phi[t] = in ? f.phi[c[t]] : 0.0f;
    }
#pragma unroll
    for (int t = 0; t < kStreamSlices; ++t) {
        const bool fr = in && freeze_on && fc[t] != GAS;
        T[t] = fr ? f.T_out[c[t]] : 0.0f;
        fl[t] = fr ? f.fl_out[c[t]] : 0.0f;
    }
#pragma unroll
    for (int t = 0; t < kStreamSlices; ++t) {
        const int b = b0 + t;
        if (b >= d.slow()) break;
        bool quiet = true;
        if (in)
            quiet = stream_rest_cell(f, s, zyx ? a : b, j, zyx ? b : a, c[t], fc[t], fp[t], T[t],
                                     phi[t], fl[t]);
        record_quiet(q, seg_index(q, (int)blockIdx.x, j, b, d), quiet, false, false);
    }
}

/* 7 over the segments kernel 2 left in (the pull ran in its own launch) */
__global__ void __launch_bounds__(256) k_step_stream_seg(waam_step_fields f, StepConsts s)
{
    const Dims &d = s.d;
    const QuietRec q = quiet_rec(f, d);
    const float T_cold = cold_limit(s.p);
    for_needed_segments(q, *q.count, [&](size_t seg) {
        int i, j, k;
        const bool in = seg_cell(q, seg, d, i, j, k);
        const size_t c = in ? d.idx(i, j, k) : 0;
        const uint8_t fc = in ? f.flags_out[c] : (uint8_t)GAS;
        const uint8_t fp = in ? f.sflags_prev[c] : (uint8_t)GAS;
        const float phi = in ? f.phi[c] : 0.0f;
        /* the freeze inputs, and the cold test's */
        const bool fr = in && fc != GAS;
        const float T = fr ? f.T_out[c] : 0.0f;
        const float fl = fr ? f.fl_out[c] : 0.0f;
        bool quiet = true;
        if (in) quiet = stream_rest_cell(f, s, i, j, k, c, fc, fp, T, phi, fl);
        /* SOLID after the flag update with phi 1, f_l 0 and T below the limit: no freeze rule
         * changes it (the cooled rule rewrites the same values), so it ends the step SOLID */
        const bool cold = !in || (fc == SOLID && phi == 1.0f && fl == 0.0f && T < T_cold);
        record_quiet(q, seg, quiet, cold, true);
    });
}

/* 7a — the previous step's pull in a launch of its own, ahead of kernel 7 (which changes
 *      flags_out): every cell this step does not collide and that was not SOLID at the step
 *      being completed pulls with that step's flags. A pure streaming pass keeps more loads in
 *      flight than the pull inside kernel 7's per-cell work. */
template <class E>
__global__ void __launch_bounds__(256) k_step_pull(waam_step_fields f, StepConsts s)
{
    const Dims &d = s.d;
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    size_t c = d.idx(i, j, k);
    uint8_t fc = f.flags_out[c];
    uint8_t fp = f.sflags_prev[c];
    bool need = !is_surface_class(fc) && fp != SOLID;
    if (__ballot(need) != 0ull)
        pull_rows(static_cast<const E *>(f.f_src), static_cast<E *>(f.f_dst), f.sflags_prev, i, j,
                  k, need, fp, d);
}

/* neighbour's position along the contiguous axis: x wraps, z clamps (reference stream,
 * kernels.py:2991-3010) */
__device__ __forceinline__ int fast_neighbour(int a, int sh, const Dims &d)
{
    int n = a - sh, nf = d.fast();
    if (d.layout == kLayoutZYX) return n < 0 ? n + nf : (n >= nf ? n - nf : n);
    return imax(0, imin(n, nf - 1));
}

/* 7a with 16-bit storage — the pull on pairs of cells along the contiguous axis: each lane
 * moves one 32-bit word (the two cells' populations) per direction, and the neighbouring pair's
 * word comes from the adjacent lane, so every load and store is a full dword of a coalesced
 * row. Lanes whose neighbour word lies outside the wavefront or across the domain edge load
 * their two values one by one. Values are those of the cell-by-cell pull. Needs an even
 * contiguous extent (the host launches k_step_pull<uint16_t> otherwise). */
__global__ void __launch_bounds__(256) k_step_pull_pair16(waam_step_fields f, StepConsts s)
{
    const Dims &d = s.d;
    const int lane = (int)threadIdx.x;
    const int nw = d.fast() >> 1;
    const int w = (int)(blockIdx.x * blockDim.x) + lane;
    const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    const int b = (int)blockIdx.z;
    if (w >= nw || j >= d.ny) return;
    const bool zyx = d.layout == kLayoutZYX;
    const int a0 = 2 * w;
    const size_t N = d.n();
    const size_t c0 = zyx ? d.idx(a0, j, b) : d.idx(b, j, a0);
    const uint16_t fc2 = *reinterpret_cast<const uint16_t *>(f.flags_out + c0);
    const uint16_t fp2 = *reinterpret_cast<const uint16_t *>(f.sflags_prev + c0);
    const uint8_t fc0 = (uint8_t)(fc2 & 0xffu), fc1 = (uint8_t)(fc2 >> 8);
    const uint8_t fp0 = (uint8_t)(fp2 & 0xffu), fp1 = (uint8_t)(fp2 >> 8);
    const bool need0 = !is_surface_class(fc0) && fp0 != SOLID;
    const bool need1 = !is_surface_class(fc1) && fp1 != SOLID;
    if (__ballot(need0 || need1) == 0ull) return;
    const uint16_t *src = static_cast<const uint16_t *>(f.f_src);
    uint16_t *dst = static_cast<uint16_t *>(f.f_dst);
    /* the row each direction pulls from (the other two axes' neighbours) */
    auto row_of = [&](int q) {
        int nj = imax(0, imin(j - ey(q), d.ny - 1));
        int nb = zyx ? imax(0, imin(b - ez(q), d.nz - 1))
                     : (b - ex(q) < 0 ? b - ex(q) + d.nx
                                      : (b - ex(q) >= d.nx ? b - ex(q) - d.nx : b - ex(q)));
        return src + (size_t)q * N + (zyx ? d.idx(0, nj, nb) : d.idx(nb, nj, 0));
    };
    /* every load first, then the shuffles: all 19 words in flight at once */
    uint32_t v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) v[q] = reinterpret_cast<const uint32_t *>(row_of(q))[w];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int sh = zyx ? ex(q) : ez(q);
        if (sh == 0) continue;
        const uint32_t aw = v[q];
        const uint32_t ow = (uint32_t)__shfl((int)aw, lane - sh);
        v[q] = sh > 0 ? ((aw << 16) | (ow >> 16)) : ((aw >> 16) | (ow << 16));
    }
    /* lanes whose neighbour word is in another wavefront or across the domain edge */
    if (lane == 0 || lane == 63 || w == 0 || w == nw - 1) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int sh = zyx ? ex(q) : ez(q);
            const bool direct = sh > 0 ? (lane == 0 || w == 0)
                                       : (sh < 0 && (lane == 63 || w == nw - 1));
            if (direct) {
                const uint16_t *row = row_of(q);
                v[q] = (uint32_t)row[fast_neighbour(a0, sh, d)] |
                       ((uint32_t)row[fast_neighbour(a0 + 1, sh, d)] << 16);
            }
        }
    }
    /* bounce-back off SOLID neighbours for a destination that was not GAS at the step whose
     * stream this is (its flag then, as every other flag of that stream) */
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const bool nh = h == 0 ? need0 : need1;
        const uint8_t fph = h == 0 ? fp0 : fp1;
        if (!nh || fph == GAS) continue;
        const int a = a0 + h;
        const int i = zyx ? a : b, k = zyx ? b : a;
        const size_t c = c0 + (size_t)h;
        for (int q = 0; q < Q; ++q) {
            int ni = i - ex(q), nj = j - ey(q), nk = k - ez(q);
            ni = ni < 0 ? ni + d.nx : (ni >= d.nx ? ni - d.nx : ni);
            nj = imax(0, imin(nj, d.ny - 1));
            nk = imax(0, imin(nk, d.nz - 1));
            if (f.sflags_prev[d.idx(ni, nj, nk)] == SOLID) {
                uint32_t r = src[(size_t)opp(q) * N + c];
                v[q] = h == 0 ? ((v[q] & 0xffff0000u) | r) : ((v[q] & 0xffffu) | (r << 16));
            }
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        uint16_t *o = dst + (size_t)q * N + c0;
        if (need0 && need1) *reinterpret_cast<uint32_t *>(o) = v[q];
        else if (need0) o[0] = (uint16_t)(v[q] & 0xffffu);
        else if (need1) o[1] = (uint16_t)(v[q] >> 16);
    }
}

/* 7a on 8-byte words — the pull with every lane moving one 64-bit word per direction: two fp32
 * cells (EB = 32) or four 16-bit cells (EB = 16) along the contiguous axis, so each of the 19
 * loads and stores of a lane is a dwordx2 of a coalesced row. The neighbouring word's edge
 * element comes from the adjacent lane (one 32-bit shuffle per direction with a contiguous-axis
 * component); lanes whose neighbour element lies in another wavefront or across the domain edge
 * load that one element directly. Values are those of the cell-by-cell pull. Needs a contiguous
 * extent divisible by 64 / EB (the host launches the narrower pulls otherwise). A block's
 * wavefronts lie side by side along the row (pull_block), so the cache lines holding the edge
 * elements a wavefront takes from beyond its own words are fetched by the same CU. Directions
 * QLO..QHI-1 only: the pull runs as two launches of 10 and 9 directions, each with half the
 * concurrent row streams (768x384x192: fp16 719 -> 710 us, fp32 1418 -> 1383 us; three or
 * four launches were slower). */
template <int EB, int QLO, int QHI>
__global__ void __launch_bounds__(256) k_step_pull_x2(waam_step_fields f, StepConsts s)
{
    constexpr int CPL = 64 / EB;  // cells per lane
    using Elem = typename std::conditional<EB == 32, uint32_t, uint16_t>::type;
    constexpr uint64_t kMask = EB == 32 ? 0xffffffffull : 0xffffull;
    const Dims &d = s.d;
    const int lane = (int)(threadIdx.x & 63u);
    const int nw = d.fast() / CPL;
    const int w = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    const int b = (int)blockIdx.z;
    if (w >= nw || j >= d.ny) return;
    const bool zyx = d.layout == kLayoutZYX;
    const int a0 = CPL * w;
    const size_t N = d.n();
    const size_t c0 = zyx ? d.idx(a0, j, b) : d.idx(b, j, a0);
    /* per cell: whether it pulls, and whether it was GAS at the step being completed (no
     * bounce-back then). The first launch reads the CPL cells' flags (one load each) and leaves
     * these bits in the lane's mask byte, which the second launch reads instead */
    uint8_t *mask = f.quiet + quiet_pull_offset(d) + c0 / CPL;
    bool need[CPL], gas[CPL];
    if (QLO == 0) {
        const uint32_t fcw = CPL == 4 ? *reinterpret_cast<const uint32_t *>(f.flags_out + c0)
                                      : *reinterpret_cast<const uint16_t *>(f.flags_out + c0);
        const uint32_t fpw = CPL == 4 ? *reinterpret_cast<const uint32_t *>(f.sflags_prev + c0)
                                      : *reinterpret_cast<const uint16_t *>(f.sflags_prev + c0);
        unsigned mb = 0;
#pragma unroll
        for (int h = 0; h < CPL; ++h) {
            const uint8_t fch = (uint8_t)((fcw >> (8 * h)) & 0xffu);
            const uint8_t fph = (uint8_t)((fpw >> (8 * h)) & 0xffu);
            need[h] = !is_surface_class(fch) && fph != SOLID;
            gas[h] = fph == GAS;
            mb |= (need[h] ? 1u : 0u) << h;
            mb |= (gas[h] ? 1u : 0u) << (4 + h);
        }
        *mask = (uint8_t)mb;
    } else {
        const unsigned mb = *mask;
#pragma unroll
        for (int h = 0; h < CPL; ++h) {
            need[h] = ((mb >> h) & 1u) != 0u;
            gas[h] = ((mb >> (4 + h)) & 1u) != 0u;
        }
    }
    bool any = false, all = true;
#pragma unroll
    for (int h = 0; h < CPL; ++h) {
        any = any || need[h];
        all = all && need[h];
    }
    if (__ballot(any) == 0ull) return;
    const Elem *src = static_cast<const Elem *>(f.f_src);
    Elem *dst = static_cast<Elem *>(f.f_dst);
    /* the row each direction pulls from (the other two axes' neighbours) */
    auto row_of = [&](int q) {
        int nj = imax(0, imin(j - ey(q), d.ny - 1));
        int nb = zyx ? imax(0, imin(b - ez(q), d.nz - 1))
                     : (b - ex(q) < 0 ? b - ex(q) + d.nx
                                      : (b - ex(q) >= d.nx ? b - ex(q) - d.nx : b - ex(q)));
        return src + (size_t)q * N + (zyx ? d.idx(0, nj, nb) : d.idx(nb, nj, 0));
    };
    /* every load first, then the shuffles: all 19 words in flight at once */
    uint64_t v[Q];
#pragma unroll
    for (int q = QLO; q < QHI; ++q) v[q] = reinterpret_cast<const uint64_t *>(row_of(q))[w];
    const bool first = lane == 0 || w == 0, last = lane == 63 || w == nw - 1;
#pragma unroll
    for (int q = QLO; q < QHI; ++q) {
        const int sh = zyx ? ex(q) : ez(q);
        if (sh == 0) continue;
        const uint64_t aw = v[q];
        if (sh > 0) {
            /* cell a0 takes the element below it: the top element of the lower word */
            const uint32_t o = (uint32_t)__shfl((int)(uint32_t)(aw >> 32), lane - 1);
            uint64_t e = (uint64_t)(o >> (32 - EB));
            if (first) e = (uint64_t)row_of(q)[fast_neighbour(a0, sh, d)];
            v[q] = (aw << EB) | e;
        } else {
            /* cell a0 + CPL - 1 takes the element above it: the bottom one of the upper word */
            const uint32_t o = (uint32_t)__shfl((int)(uint32_t)aw, lane + 1);
            uint64_t e = (uint64_t)o & kMask;
            if (last) e = (uint64_t)row_of(q)[fast_neighbour(a0 + CPL - 1, sh, d)];
            v[q] = (aw >> EB) | (e << (64 - EB));
        }
    }
    /* bounce-back off SOLID neighbours for a destination that was not GAS at the step whose
     * stream this is (its flag then, as every other flag of that stream) */
#pragma unroll
    for (int h = 0; h < CPL; ++h) {
        if (!need[h] || gas[h]) continue;
        const int a = a0 + h;
        const int i = zyx ? a : b, k = zyx ? b : a;
        const size_t c = c0 + (size_t)h;
        for (int q = QLO; q < QHI; ++q) {
            int ni = i - ex(q), nj = j - ey(q), nk = k - ez(q);
            ni = ni < 0 ? ni + d.nx : (ni >= d.nx ? ni - d.nx : ni);
            nj = imax(0, imin(nj, d.ny - 1));
            nk = imax(0, imin(nk, d.nz - 1));
            if (f.sflags_prev[d.idx(ni, nj, nk)] == SOLID) {
                const uint64_t r = (uint64_t)src[(size_t)opp(q) * N + c];
                v[q] = (v[q] & ~(kMask << (EB * h))) | (r << (EB * h));
            }
        }
    }
#pragma unroll
    for (int q = QLO; q < QHI; ++q) {
        Elem *o = dst + (size_t)q * N + c0;
        if (all) {
            /* a non-temporal store: the next step reads these words, nothing in this one
             * (1-2 % faster than a plain store; non-temporal loads were slower) */
            __builtin_nontemporal_store(v[q], reinterpret_cast<uint64_t *>(o));
        } else {
#pragma unroll
            for (int h = 0; h < CPL; ++h)
                if (need[h]) o[h] = (Elem)((v[q] >> (EB * h)) & kMask);
        }
    }
}

/* launch shape of k_step_pull_x2 with `cpl` cells per lane: m wavefronts side by side along a
 * row per block, m the largest divisor of the row's wavefront count up to 4 (rows of 4 / m
 * per block when m divides 4) */
static void pull_block(const Dims &d, int cpl, dim3 &grid, dim3 &block)
{
    const int waves = (d.fast() / cpl + kTileFast - 1) / kTileFast;
    int m = 4;
    while (waves % m != 0) --m;
    const int rows = 4 % m == 0 ? 4 / m : 1;
    block = dim3((unsigned)(kTileFast * m), (unsigned)rows, 1);
    grid = dim3((unsigned)(waves / m), (unsigned)((d.ny + rows - 1) / rows), (unsigned)d.slow());
}

/* standalone pending stream */
template <class E>
__global__ void k_stream_pending(const E *__restrict__ src, E *__restrict__ dst,
                                 const uint8_t *__restrict__ sfl, Dims d)
{
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    uint8_t fp = sfl[d.idx(i, j, k)];
    bool need = fp != SOLID;
    if (__ballot(need) != 0ull) pull_rows(src, dst, sfl, i, j, k, need, fp, d);
}

/* 6 — advect_tracers (reference kernels.py:3206-3255) */
__global__ void k_advect_tracers_step(float *__restrict__ pos, int32_t *__restrict__ active,
                                      const float *__restrict__ ux, const float *__restrict__ uy,
                                      const float *__restrict__ uz, const float *__restrict__ f_l,
                                      const uint8_t *__restrict__ flags, float dx, int max_tracers,
                                      Dims d)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= max_tracers || active[p] != 1) return;
    float px = pos[3 * p] / dx, py = pos[3 * p + 1] / dx, pz = pos[3 * p + 2] / dx;
    int i = (int)px, j = (int)py, k = (int)pz;
    if (i < 0 || i >= d.nx || j < 0 || j >= d.ny || k < 0 || k >= d.nz) {
        active[p] = 0;
        return;
    }
    size_t c = d.idx(i, j, k);
    uint8_t fg = flags[c];
    if (fg == GAS) {
        active[p] = 0;
        return;
    }
    if (f_l[c] < 0.05f || fg == SOLID) {
        active[p] = 2;
        return;
    }
    pos[3 * p] += ux[c] * dx;
    pos[3 * p + 1] += uy[c] * dx;
    pos[3 * p + 2] += uz[c] * dx;
}

/* half-widths (cells) of the (i, j) box outside which the arc heat is exactly zero: exp(x) is
 * exactly zero for x < -103.98, and outside the box x < -110 for every cell; r < 0 when the
 * shape has no finite box */
static void arc_radii(const StepConsts &s, double &ri, double &rj)
{
    const waam_step_params &p = s.p;
    if (p.arc_kind == 0) {
        double inv = (double)s.arc.inv2s2;
        ri = rj = std::sqrt(110.0 / inv);
    } else {
        double a = std::max((double)p.g_af, (double)p.g_ar);
        ri = std::sqrt(110.0 / 3.0 * (a * a + 1e-6)) / std::max(std::fabs((double)p.g_sgn), 1e-30);
        rj = std::sqrt(110.0 / 3.0 * ((double)p.g_b * p.g_b + 1e-6));
    }
    if (!(ri < 1e6) || !(rj < 1e6)) ri = rj = -1.0;
}

/* the arc box of a centre (arc_i, arc_j), as the kernels compare it with cell indices */
static void arc_box(double ri, double rj, float arc_i, float arc_j, ArcStep &a)
{
    if (ri < 0.0) {
        a.box_i0 = a.box_j0 = -(1 << 30);
        a.box_i1 = a.box_j1 = (1 << 30);
        return;
    }
    auto lo = [](double c, double r) { return (int)std::floor(c - r) - 2; };
    auto hi = [](double c, double r) { return (int)std::ceil(c + r) + 2; };
    a.box_i0 = lo(arc_i, ri);
    a.box_i1 = hi(arc_i, ri);
    a.box_j0 = lo(arc_j, rj);
    a.box_j1 = hi(arc_j, rj);
}

/* HIP-event timing of one kernel kind (waam_timing_begin / waam_timing_end) */
struct KernelTimer {
    bool armed = false;
    int which = -1, used = 0;
    int every = 1;
    int64_t calls = 0;           // launches of `which` since arming
    std::vector<hipEvent_t> ev;  // start/stop pairs
};
static KernelTimer g_timer;

/* launch(start, stop) launches the kernel(s) of one kind with hipExtLaunchKernelGGL, which
 * records the two events with the dispatch itself (null events when not timing) */
template <class Launch>
static void timed_launch(int which, Launch &&launch)
{
    KernelTimer &t = g_timer;
    bool rec = false;
    if (t.armed && t.which == which)
        rec = (t.calls++ % t.every) == 0 && 2 * t.used < (int)t.ev.size();
    launch(rec ? t.ev[2 * t.used] : nullptr, rec ? t.ev[2 * t.used + 1] : nullptr);
    if (rec) ++t.used;
}

/* the checks waam_step_fused makes on a field set */
static int check_fields(const char *who, const waam_step_fields &f)
{
    WAAM_RETURN_IF(check_dims(who, f.nx, f.ny, f.nz, f.layout));
    WAAM_RETURN_IF(check_ptrs(who, {f.f_src, f.f_dst, f.rho, f.ux, f.uy, f.uz, f.H_in, f.H_out,
                                    f.T_in, f.T_out, f.fl_in, f.fl_out, f.phi, f.phi_tmp,
                                    f.flags_in, f.flags_out, f.sflags_prev, f.sflags_next, f.Fx,
                                    f.Fy, f.Fz, f.T_max, f.T_prev, f.dT_dt, f.t800, f.t1100, f.tsol,
                                    f.list_thermal, f.list_collide, f.scratch, f.quiet}));
    if (f.f_src == f.f_dst || f.H_in == f.H_out || f.T_in == f.T_out || f.fl_in == f.fl_out ||
        f.flags_in == f.flags_out || f.sflags_prev == f.sflags_next || f.phi == f.phi_tmp ||
        f.list_thermal == f.list_collide)
        return set_error(WAAM_ERR_INVALID, "%s: an in/out buffer pair aliases", who);
    Dims d{f.nx, f.ny, f.nz, f.layout};
    if (d.n() > (size_t)INT_MAX)
        return set_error(WAAM_ERR_INVALID, "%s: %zu cells exceed the int32 cell lists", who, d.n());
    if (f.max_tracers > 0 && (!f.tracer_pos || !f.tracer_active))
        return set_error(WAAM_ERR_INVALID, "%s: tracer buffers missing", who);
    if (f.pop_dtype != WAAM_POP_F32 && f.pop_dtype != WAAM_POP_F16 && f.pop_dtype != WAAM_POP_BF16)
        return set_error(WAAM_ERR_INVALID, "%s: unknown population form %d", who, f.pop_dtype);
    return 0;
}

/* the temperature-dependent material's fields, when a step uses them */
static int check_tables(const char *who, const waam_step_params &p, const waam_step_fields &f)
{
    if (p.collision < 0 || p.collision > 2)
        return set_error(WAAM_ERR_INVALID, "%s: unknown collision %d", who, p.collision);
    if (p.collision == 2 && !p.use_tables)
        return set_error(WAAM_ERR_INVALID, "%s: tau per cell needs use_tables", who);
    if (!p.use_tables) return 0;
    WAAM_RETURN_IF(check_ptrs(who, {f.cp_rho_prev, f.cp_rho, f.alpha_lu, f.dgamma_lu, f.tau}));
    if (f.cp_rho_prev == f.cp_rho)
        return set_error(WAAM_ERR_INVALID, "%s: cp_rho_prev aliases cp_rho", who);
    for (int n : {p.tables.n_cp, p.tables.n_k, p.tables.n_mu, p.tables.n_dgamma})
        if (n < 0 || n > WAAM_MAX_KNOTS)
            return set_error(WAAM_ERR_INVALID, "%s: a material table holds %d knots", who, n);
    return 0;
}

/* kernel arguments of a step: the scalar groups each kernel reads, derived once on the host */
static StepConsts make_step_consts(const waam_step_params &p, const Dims &d)
{
    StepConsts s{};
    s.p = p;
    s.d = d;
    /* the quiet record is continued only on steady steps (see waam_step_params.quiet_skip) */
    s.p.quiet_skip = (p.quiet_skip && p.pending_stream == 1 && p.thermal_list_ready != 0 &&
                      !p.arc_applied && !p.forces_from_fields && !p.force_reset_all)
                         ? 1
                         : 0;
    s.arc.kind = p.arc_kind;
    s.arc.arc_i = p.arc_i;
    s.arc.arc_j = p.arc_j;
    s.arc.inv2s2 = 1.0f / (2.0f * p.sigma * p.sigma);
    s.arc.sgn = p.g_sgn;
    s.arc.ff = p.g_ff;
    s.arc.fr = p.g_fr;
    s.arc.af = p.g_af;
    s.arc.ar = p.g_ar;
    s.arc.b = p.g_b;
    s.arc.cax = p.g_c;
    s.arc.pen = p.pen;
    s.arc.enable = p.surface_weight;
    s.pc = make_phase_consts(p.cp_rho, p.L_rho, p.T_s, p.T_l);
    if (p.use_tables) {
        /* update_phase_variable_cp takes the material's H_sol and H_liq (reference
         * coupled_step.py:228-233); the cell's cp*rho comes per cell */
        s.pc.H_sol = p.H_sol_tab;
        s.pc.H_liq = p.H_liq_tab;
    }
    s.lc = LossConsts{p.T_amb, p.h_conv, p.eps_rad, p.sigma_sb, p.dt, p.dx, p.enable_conv, p.enable_rad};
    s.cap_l = s.pc.H_liq + p.cp_rho * (p.T_cap - p.T_l);
    s.cap_s = p.cp_rho * p.T_cap;
    s.h_min = p.cp_rho * p.T_amb;
    s.inv_dt = 1.0f / (p.dt + 1e-12f);
    s.ap.arc_i = p.arc_i;
    s.ap.arc_j = p.arc_j;
    s.ap.inv2s2 = 1.0f / (2.0f * p.p_sigma * p.p_sigma + 1e-6f);
    s.ap.F_peak_lu = p.p_pa / (p.rho_phys * p.dx) * (p.dt * p.dt) / p.dx;
    /* recoil and gas shear around the arc axis with the arc's sigma (reference
     * physics/weld_forces.py:128-161) */
    s.rc = Recoil{};
    s.rc.arc_i = p.arc_i;
    s.rc.arc_j = p.arc_j;
    s.rc.inv2s2 = 1.0f / (2.0f * p.sigma * p.sigma + 1e-6f);
    s.rc.clausius = p.recoil_clausius;
    s.rc.F_peak_lu = p.recoil_pa / (p.rho_phys * p.dx) * p.dt * p.dt / p.dx;
    s.rc.T_ref = p.recoil_T_ref;
    s.rc.P_ref = p.P_vap_ref;
    s.rc.T_boil = p.T_boil;
    s.rc.L_vap = p.L_vap;
    s.rc.R_spec = p.R_vap;
    s.rc.C_acc = p.C_acc;
    s.rc.dt = p.dt;
    s.rc.dx = p.dx;
    s.rc.rho_ref = p.rho_phys;
    s.gs.arc_i = p.arc_i;
    s.gs.arc_j = p.arc_j;
    s.gs.inv2s2 = s.rc.inv2s2;
    s.gs.tau_peak = p.gas_tau;
    s.gs.F_scale = 1.0f / (p.rho_phys * p.dx) * p.dt * p.dt / p.dx;
    s.at.arc_i = p.arc_i;
    s.at.arc_j = p.arc_j;
    s.at.k_robot = p.k_robot;
    s.at.col_i = p.col_i;
    s.at.col_j = p.col_j;
    s.at.dir_x = p.dir_x;
    s.at.dir_y = p.dir_y;
    s.at.dir_z = p.dir_z;
    double ri, rj;
    arc_radii(s, ri, rj);
    arc_box(ri, rj, p.arc_i, p.arc_j, s.at);
    s.track = nullptr;
    return s;
}

/* one launch: timed with events (hipExtLaunchKernel) in a direct step, a plain launch into a
 * stream being captured */
template <class T> struct ident { using type = T; };

template <typename... KArgs>
static void launch_k(bool capture, void (*k)(KArgs...), dim3 grid, dim3 block, hipStream_t st,
                     hipEvent_t e0, hipEvent_t e1, typename ident<KArgs>::type... args)
{
    if (capture) {
        void *argv[] = {static_cast<void *>(&args)...};
        (void)hipLaunchKernel(reinterpret_cast<const void *>(k), grid, block, argv, 0, st);
    } else {
        hipExtLaunchKernelGGL(k, grid, block, 0, st, e0, e1, 0, args...);
    }
}

/* the step's launches in order (kernels 1-7). A steady step (s.p.quiet_skip) lists its active
 * row segments after kernel 1 and runs kernels 2, 4 and 7 over that list only, with the surface
 * cells of kernel 3 and the pull in launches of their own. */
static void launch_step(const StepConsts &s, const waam_step_fields &f, hipStream_t st,
                        bool capture)
{
    const waam_step_params &p = s.p;
    const Dims &d = s.d;
    dim3 g3 = grid3(d);
    const bool seg = p.quiet_skip != 0;
    int do_sums = p.is_welding && !p.arc_applied ? 1 : 0;
    timed_launch(WAAM_K_BEGIN, [&](hipEvent_t a, hipEvent_t b) {
        launch_k(capture, k_step_begin, dim3(do_sums ? kReduceBlocks : 1), dim3(kBlock), st, a,
                 (do_sums || seg) ? nullptr : b, f, s, do_sums);
        if (do_sums)
            launch_k(capture, k_step_arc_total, dim3(1), dim3(kBlock), st, nullptr,
                     seg ? nullptr : b, f, s);
        if (seg)
            launch_k(capture, k_step_quiet_list, dim3(grid_for(quiet_segments(d))), dim3(kBlock),
                     st, nullptr, b, f, s);
    });
    /* kernel 3 (the listed surface cells), with or without the temperature-dependent material */
    const bool tab = p.use_tables != 0;
    auto launch_surface = [&](hipEvent_t a, hipEvent_t b) {
        if (tab)
            launch_k(capture, k_step_thermal_surface<true>,
                     dim3(resident_blocks(k_step_thermal_surface<true>, kBlock)), dim3(kBlock), st,
                     a, b, f, s);
        else
            launch_k(capture, k_step_thermal_surface<false>,
                     dim3(resident_blocks(k_step_thermal_surface<false>, kBlock)), dim3(kBlock),
                     st, a, b, f, s);
    };
    if (seg) {
        timed_launch(WAAM_K_THERMAL, [&](hipEvent_t a, hipEvent_t b) {
            const bool plate = plate_slices(p, d) > 0;
            if (plate)
                launch_k(capture, k_step_thermal_plate,
                         dim3((unsigned)quiet_nbx(d), (unsigned)((d.ny + kPlateRows - 1) / kPlateRows),
                              (unsigned)kPlateChunks),
                         dim3(kTileFast, kPlateRows), st, a, nullptr, f, s);
            if (tab)
                launch_k(capture, k_step_thermal_seg<true>,
                         dim3(resident_blocks(k_step_thermal_seg<true>, kBlock)), block3(), st,
                         plate ? nullptr : a, b, f, s);
            else
                launch_k(capture, k_step_thermal_seg<false>,
                         dim3(resident_blocks(k_step_thermal_seg<false>, kBlock)), block3(), st,
                         plate ? nullptr : a, b, f, s);
        });
        timed_launch(WAAM_K_THERMAL_SURFACE, launch_surface);
    } else if (p.thermal_list_ready == 1) {
        /* enough list slices for about kListBlocksWithSweep workers */
        int per = (int)(g3.x * g3.y);
        int lz = std::max(1, (kListBlocksWithSweep + per - 1) / per);
        dim3 ga(g3.x, g3.y, g3.z + (unsigned)lz);
        timed_launch(WAAM_K_THERMAL, [&](hipEvent_t a, hipEvent_t b) {
            if (tab) launch_k(capture, k_step_thermal_all<true>, ga, block3(), st, a, b, f, s, lz);
            else launch_k(capture, k_step_thermal_all<false>, ga, block3(), st, a, b, f, s, lz);
        });
    } else {
        timed_launch(WAAM_K_THERMAL, [&](hipEvent_t a, hipEvent_t b) {
            if (p.thermal_list_ready) {
                if (tab) launch_k(capture, k_step_thermal<false, true>, g3, block3(), st, a, b, f, s);
                else launch_k(capture, k_step_thermal<false, false>, g3, block3(), st, a, b, f, s);
            } else {
                if (tab) launch_k(capture, k_step_thermal<true, true>, g3, block3(), st, a, b, f, s);
                else launch_k(capture, k_step_thermal<true, false>, g3, block3(), st, a, b, f, s);
            }
        });
        timed_launch(WAAM_K_THERMAL_SURFACE, launch_surface);
    }
    timed_launch(WAAM_K_VOF_FLAGS, [&](hipEvent_t a, hipEvent_t b) {
        if (seg)
            launch_k(capture, k_step_vof_flags_seg,
                     dim3(resident_blocks(k_step_vof_flags_seg, kBlock)), block3(), st, a, b, f, s);
        else launch_k(capture, k_step_vof_flags, g3, block3(), st, a, b, f, s);
    });
    timed_launch(WAAM_K_COLLIDE, [&](hipEvent_t a, hipEvent_t b) {
        if (f.pop_dtype == WAAM_POP_F16)
            launch_k(capture, k_step_collide<WAAM_POP_F16>,
                     dim3(resident_blocks(k_step_collide<WAAM_POP_F16>, kBlock)), dim3(kBlock), st,
                     a, b, f, s);
        else if (f.pop_dtype == WAAM_POP_BF16)
            launch_k(capture, k_step_collide<WAAM_POP_BF16>,
                     dim3(resident_blocks(k_step_collide<WAAM_POP_BF16>, kBlock)), dim3(kBlock), st,
                     a, b, f, s);
        else
            launch_k(capture, k_step_collide<WAAM_POP_F32>,
                     dim3(resident_blocks(k_step_collide<WAAM_POP_F32>, kBlock)), dim3(kBlock), st,
                     a, b, f, s);
    });
    if (f.max_tracers > 0)
        timed_launch(WAAM_K_TRACERS, [&](hipEvent_t a, hipEvent_t b) {
            launch_k(capture, k_advect_tracers_step, dim3(grid_for((size_t)f.max_tracers)),
                     dim3(kBlock), st, a, b, f.tracer_pos, f.tracer_active, f.ux, f.uy, f.uz,
                     f.fl_out, f.flags_out, p.dx, f.max_tracers, d);
        });
    /* the pull runs in a launch of its own ahead of kernel 7 (8-byte words, or pairs of cells
     * for 16-bit storage on an even contiguous extent), or inside kernel 7 */
    const bool f32 = f.pop_dtype == WAAM_POP_F32;
    const bool pair16 = !f32 && d.fast() % 2 == 0;
    /* 8-byte words per lane: two fp32 cells, or four 16-bit cells (16-byte words measured no
     * faster: 727 against 718 us for the 16-bit pull of 768x384x192) */
    const bool wide = kWidePull && (f32 ? d.fast() % 2 == 0 : d.fast() % 4 == 0);
    const bool separate = seg || pair16 || (f32 ? kSeparatePull32 : kSeparatePull16);
    if (separate && p.pending_stream)
        timed_launch(WAAM_K_PULL, [&](hipEvent_t a, hipEvent_t b) {
            if (wide) {
                dim3 gp, bp;
                pull_block(d, f32 ? 2 : 4, gp, bp);
                if (f32) {
                    launch_k(capture, k_step_pull_x2<32, 0, 10>, gp, bp, st, a, nullptr, f, s);
                    launch_k(capture, k_step_pull_x2<32, 10, Q>, gp, bp, st, nullptr, b, f, s);
                } else {
                    launch_k(capture, k_step_pull_x2<16, 0, 10>, gp, bp, st, a, nullptr, f, s);
                    launch_k(capture, k_step_pull_x2<16, 10, Q>, gp, bp, st, nullptr, b, f, s);
                }
            } else if (pair16) {
                dim3 gp((unsigned)((d.fast() / 2 + kTileFast - 1) / kTileFast), g3.y, g3.z);
                launch_k(capture, k_step_pull_pair16, gp, block3(), st, a, b, f, s);
            } else if (f32) {
                launch_k(capture, k_step_pull<float>, g3, block3(), st, a, b, f, s);
            } else {
                launch_k(capture, k_step_pull<uint16_t>, g3, block3(), st, a, b, f, s);
            }
        });
    timed_launch(WAAM_K_STREAM, [&](hipEvent_t a, hipEvent_t b) {
        dim3 gs(g3.x, g3.y, (unsigned)((d.slow() + kStreamSlices - 1) / kStreamSlices));
        if (seg)
            launch_k(capture, k_step_stream_seg, dim3(resident_blocks(k_step_stream_seg, kBlock)),
                     block3(), st, a, b, f, s);
        else if (separate)
            launch_k(capture, k_step_stream<float, false>, gs, block3(), st, a, b, f, s);
        else if (f32)
            launch_k(capture, k_step_stream<float, true>, g3, block3(), st, a, b, f, s);
        else
            launch_k(capture, k_step_stream<uint16_t, true>, g3, block3(), st, a, b, f, s);
    });
}

}  // namespace waam

/* a captured run of fused steps (waam_graph_capture) */
struct waam_step_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int n_steps = 0, device = 0;
    int nx = 0, ny = 0;
    waam::StepConsts base{};            // the captured steps' scalars (arc of step 0)
    double ri = 0.0, rj = 0.0;          // arc box half-widths
    waam::ArcStep *track = nullptr;     // device: one entry per captured step
    waam::ArcStep *stage[2] = {nullptr, nullptr};  // pinned host staging, used in turn
    hipEvent_t staged[2] = {nullptr, nullptr};     // the copy out of each staging buffer
    int next = 0;
};

using namespace waam;

static void graph_free(waam_step_graph *g)
{
    if (!g) return;
    /* a replay may still be running or its track copy pending: let the device drain before the
     * graph, its track and the staging buffers go */
    if (g->exec) (void)hipDeviceSynchronize();
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    for (int b = 0; b < 2; ++b) {
        if (g->staged[b]) {
            (void)hipEventSynchronize(g->staged[b]);
            (void)hipEventDestroy(g->staged[b]);
        }
        if (g->stage[b]) (void)hipHostFree(g->stage[b]);
    }
    if (g->track) (void)hipFree(g->track);
    delete g;
}

namespace waam {
static void launch_droplet_step(const StepConsts &s0, const waam_step_fields &f,
                                const waam_droplet &dp, hipStream_t st, bool capture);
static int check_droplet(const char *who, const waam_step_fields &f, const waam_droplet &dp);
}  // namespace waam

extern "C" {

int waam_step_fused(const waam_step_params *pp, const waam_step_fields *ff, void *stream)
{
    if (!pp || !ff) return set_error(WAAM_ERR_INVALID, "waam_step_fused: NULL argument");
    const waam_step_params &p = *pp;
    const waam_step_fields &f = *ff;
    WAAM_RETURN_IF(check_fields("waam_step_fused", f));
    WAAM_RETURN_IF(check_tables("waam_step_fused", p, f));
    if (p.col_i < 0 || p.col_i >= f.nx || p.col_j < 0 || p.col_j >= f.ny)
        return set_error(WAAM_ERR_INVALID, "waam_step_fused: arc column outside the grid");
    Dims d{f.nx, f.ny, f.nz, f.layout};
    StepConsts s = make_step_consts(p, d);
    launch_step(s, f, as_stream(stream), false);
    return check_launch("waam_step_fused");
}

}  // extern "C"

namespace waam {

/* Allocation-free step: the workspace and the two cell lists of waam_step_fused_noalloc live in
 * this file-scope device array (each GPU that runs the entry holds its own copy, with its own
 * address, until the process ends): the workspace first, then the thermal list, then the
 * collision list. */
constexpr size_t kNoAllocMaxCells = (size_t)1 << 22;
constexpr size_t kNoAllocListOffset = (kScratchBytes + 255) / 256 * 256;
constexpr size_t kNoAllocBytes = kNoAllocListOffset + 2 * kNoAllocMaxCells * sizeof(int32_t);
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[kNoAllocMaxDevices];

/* the device address of `scratch` on the current device, looked up once per device */
static int noalloc_scratch(void **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return set_error(WAAM_ERR_HIP, "waam_step_fused_noalloc: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        return set_error(WAAM_ERR_INVALID, "waam_step_fused_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(WAAM_ERR_HIP, "waam_step_fused_noalloc: device array lookup: %s",
                             hipGetErrorString(e));
        g_noalloc_ptr[dev] = ptr;
    }
    *out = g_noalloc_ptr[dev];
    return WAAM_OK;
}

}  // namespace waam

extern "C" {

int64_t waam_noalloc_max_cells(void) { return (int64_t)kNoAllocMaxCells; }

int64_t waam_quiet_bytes(int nx, int ny, int nz, int layout)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || (layout != kLayoutXYZ && layout != kLayoutZYX)) return 0;
    return (int64_t)quiet_bytes(Dims{nx, ny, nz, layout});
}

int waam_step_fused_noalloc(const waam_step_params *pp, const waam_step_fields *ff, void *stream)
{
    if (!pp || !ff) return set_error(WAAM_ERR_INVALID, "waam_step_fused_noalloc: NULL argument");
    const waam_step_params &p = *pp;
    const int64_t n = (int64_t)ff->nx * ff->ny * ff->nz;
    if (ff->nx <= 0 || ff->ny <= 0 || ff->nz <= 0 || n > (int64_t)kNoAllocMaxCells)
        return set_error(WAAM_ERR_INVALID,
                         "waam_step_fused_noalloc: grid %d x %d x %d (%lld cells) outside 1..%lld "
                         "cells, the size of its device work array; use waam_step_fused",
                         ff->nx, ff->ny, ff->nz, (long long)n, (long long)kNoAllocMaxCells);
    void *base = nullptr;
    WAAM_RETURN_IF(noalloc_scratch(&base));
    waam_step_fields f = *ff;
    f.scratch = base;
    f.list_thermal = reinterpret_cast<int32_t *>((char *)base + kNoAllocListOffset);
    f.list_collide = f.list_thermal + kNoAllocMaxCells;
    WAAM_RETURN_IF(check_fields("waam_step_fused_noalloc", f));
    WAAM_RETURN_IF(check_tables("waam_step_fused_noalloc", p, f));
    if (p.col_i < 0 || p.col_i >= f.nx || p.col_j < 0 || p.col_j >= f.ny)
        return set_error(WAAM_ERR_INVALID, "waam_step_fused_noalloc: arc column outside the grid");
    Dims d{f.nx, f.ny, f.nz, f.layout};
    StepConsts s = make_step_consts(p, d);
    launch_step(s, f, as_stream(stream), false);
    return check_launch("waam_step_fused_noalloc");
}

int waam_graph_capture(const waam_step_params *pp, const waam_step_fields *f_even,
                       const waam_step_fields *f_odd, int n_steps, waam_step_graph **out)
{
    return waam_graph_capture_plan(pp, f_even, f_odd, n_steps, nullptr, nullptr, out);
}

int waam_graph_capture_plan(const waam_step_params *pp, const waam_step_fields *f_even,
                            const waam_step_fields *f_odd, int n_steps, const int32_t *kinds,
                            const waam_droplet *drops, waam_step_graph **out)
{
    if (!pp || !f_even || !f_odd || !out)
        return set_error(WAAM_ERR_INVALID, "waam_graph_capture: NULL argument");
    if (kinds && !drops)
        return set_error(WAAM_ERR_INVALID, "waam_graph_capture: a plan without droplets");
    *out = nullptr;
    const waam_step_params &p = *pp;
    if (n_steps < 1 || n_steps > kGraphMaxSteps)
        return set_error(WAAM_ERR_INVALID, "waam_graph_capture: %d steps (1..%d)", n_steps,
                         kGraphMaxSteps);
    /* a captured step is an ordinary fused step that follows a fused step: its population
     * stream is pending, its thermal list was built by that step, nothing was applied eagerly */
    if (p.pending_stream != 1 || (p.thermal_list_ready != 1 && p.thermal_list_ready != 2) ||
        p.arc_applied || p.use_arc_k_override || p.force_reset_all || p.forces_from_fields)
        return set_error(WAAM_ERR_INVALID,
                         "waam_graph_capture: only steady steps are captured (pending stream, "
                         "thermal list ready, no eager arc, no force reset)");
    if (g_timer.armed)
        return set_error(WAAM_ERR_INVALID, "waam_graph_capture: kernel timing is armed");
    WAAM_RETURN_IF(check_fields("waam_graph_capture", *f_even));
    WAAM_RETURN_IF(check_fields("waam_graph_capture", *f_odd));
    WAAM_RETURN_IF(check_tables("waam_graph_capture", p, *f_even));
    WAAM_RETURN_IF(check_tables("waam_graph_capture", p, *f_odd));
    const waam_step_fields &a = *f_even, &b = *f_odd;
    if (a.nx != b.nx || a.ny != b.ny || a.nz != b.nz || a.layout != b.layout ||
        a.pop_dtype != b.pop_dtype || a.max_tracers != b.max_tracers)
        return set_error(WAAM_ERR_INVALID, "waam_graph_capture: the two field sets differ in shape");
    /* the odd set is the even set after one step's swaps */
    if (b.f_src != a.f_dst || b.f_dst != a.f_src || b.H_in != a.H_out || b.H_out != a.H_in ||
        b.T_in != a.T_out || b.T_out != a.T_in || b.fl_in != a.fl_out || b.fl_out != a.fl_in ||
        b.flags_in != a.flags_out || b.flags_out != a.flags_in ||
        b.sflags_prev != a.sflags_next || b.sflags_next != a.sflags_prev || b.phi != a.phi ||
        b.phi_tmp != a.phi_tmp || b.list_thermal != a.list_thermal ||
        b.list_collide != a.list_collide || b.scratch != a.scratch || b.quiet != a.quiet ||
        b.cp_rho_prev != a.cp_rho || b.cp_rho != a.cp_rho_prev || b.alpha_lu != a.alpha_lu ||
        b.dgamma_lu != a.dgamma_lu || b.tau != a.tau)
        return set_error(WAAM_ERR_INVALID,
                         "waam_graph_capture: the odd field set is not the even set swapped");
    if (p.col_i < 0 || p.col_i >= a.nx || p.col_j < 0 || p.col_j >= a.ny)
        return set_error(WAAM_ERR_INVALID, "waam_graph_capture: arc column outside the grid");
    for (int k = 0; kinds && k < n_steps; ++k) {
        if (kinds[k] < 0 || kinds[k] > 2)
            return set_error(WAAM_ERR_INVALID, "waam_graph_capture: step %d of kind %d", k,
                             kinds[k]);
        if (kinds[k] == 1) {
            if (!p.is_welding || p.arc_kind > 1)
                return set_error(WAAM_ERR_INVALID,
                                 "waam_graph_capture: droplet steps need a welding step with the "
                                 "Gaussian or Goldak source");
            WAAM_RETURN_IF(check_droplet("waam_graph_capture", (k & 1) ? b : a, drops[k]));
        }
    }

    waam_step_graph *g = new waam_step_graph();
    Dims d{a.nx, a.ny, a.nz, a.layout};
    g->n_steps = n_steps;
    g->nx = a.nx;
    g->ny = a.ny;
    g->base = make_step_consts(p, d);
    arc_radii(g->base, g->ri, g->rj);
    hipStream_t cs = nullptr;
    hipError_t e = hipGetDevice(&g->device);
    const size_t tb = sizeof(ArcStep) * (size_t)n_steps;
    if (e == hipSuccess) e = hipMalloc(&g->track, tb);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = hipHostMalloc(&g->stage[k], tb, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&g->staged[k], hipEventDisableTiming);
    }
    if (e == hipSuccess) {
        /* every captured step reads the arc from the track until the first launch fills it:
         * start it as step 0 repeated */
        for (int k = 0; k < n_steps; ++k) g->stage[0][k] = g->base.at;
        e = hipMemcpy(g->track, g->stage[0], tb, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        for (int k = 0; k < n_steps; ++k) {
            StepConsts s = g->base;
            s.track = g->track + k;
            const int kind = kinds ? kinds[k] : 0;
            if (kind == 1) {
                launch_droplet_step(s, (k & 1) ? b : a, drops[k], cs, true);
            } else {
                if (kind == 2) {
                    /* the droplet step took its forces from the fields: not a steady step */
                    s.p.force_reset_all = 1;
                    s.p.quiet_skip = 0;
                }
                launch_step(s, (k & 1) ? b : a, cs, true);
            }
        }
        hipError_t le = hipGetLastError();
        e = hipStreamEndCapture(cs, &g->graph);
        if (e == hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (cs) (void)hipStreamDestroy(cs);
    if (e != hipSuccess) {
        graph_free(g);
        return set_error(WAAM_ERR_HIP, "waam_graph_capture: %s", hipGetErrorString(e));
    }
    *out = g;
    return 0;
}

int waam_graph_launch(waam_step_graph *g, const double *arc_i, const double *arc_j,
                      const float *k_robot, const float *dir, void *stream)
{
    if (!g || !arc_i || !arc_j) return set_error(WAAM_ERR_INVALID, "waam_graph_launch: NULL argument");
    const int n = g->n_steps;
    const int b = g->next;
    /* the copy that last read this staging buffer has run before it is refilled */
    hipError_t e = hipEventSynchronize(g->staged[b]);
    if (e != hipSuccess)
        return set_error(WAAM_ERR_HIP, "waam_graph_launch: %s", hipGetErrorString(e));
    ArcStep *t = g->stage[b];
    for (int k = 0; k < n; ++k) {
        const double x = arc_i[k], y = arc_j[k];
        if (!(x > -1e9 && x < 1e9) || !(y > -1e9 && y < 1e9))
            return set_error(WAAM_ERR_INVALID, "waam_graph_launch: arc centre of step %d", k);
        ArcStep a = g->base.at;
        /* the column as the host computes it for a direct step: truncation of the double
         * centre, clamped to the grid */
        a.col_i = std::max(0, std::min((int)x, g->nx - 1));
        a.col_j = std::max(0, std::min((int)y, g->ny - 1));
        a.arc_i = (float)x;
        a.arc_j = (float)y;
        if (k_robot) a.k_robot = k_robot[k];
        if (dir) {
            a.dir_x = dir[3 * k];
            a.dir_y = dir[3 * k + 1];
            a.dir_z = dir[3 * k + 2];
        }
        arc_box(g->ri, g->rj, a.arc_i, a.arc_j, a);
        t[k] = a;
    }
    hipStream_t st = as_stream(stream);
    e = hipMemcpyAsync(g->track, t, sizeof(ArcStep) * (size_t)n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(g->staged[b], st);
    if (e == hipSuccess) e = hipGraphLaunch(g->exec, st);
    if (e != hipSuccess)
        return set_error(WAAM_ERR_HIP, "waam_graph_launch: %s", hipGetErrorString(e));
    g->next = b ^ 1;
    return 0;
}

int waam_graph_steps(const waam_step_graph *g) { return g ? g->n_steps : 0; }

int waam_graph_destroy(waam_step_graph *g)
{
    graph_free(g);
    return 0;
}


int waam_stream_pending(const void *f_src, void *f_dst, const uint8_t *sflags, int nx, int ny,
                        int nz, int layout, int pop_dtype, void *stream)
{
    WAAM_RETURN_IF(check_dims("waam_stream_pending", nx, ny, nz, layout));
    WAAM_RETURN_IF(check_ptrs("waam_stream_pending", {f_src, f_dst, sflags}));
    if (f_src == f_dst) return set_error(WAAM_ERR_INVALID, "waam_stream_pending: f_src aliases f_dst");
    Dims d{nx, ny, nz, layout};
    if (pop_dtype == WAAM_POP_F32)
        hipLaunchKernelGGL(k_stream_pending<float>, grid3(d), block3(), 0, as_stream(stream),
                           static_cast<const float *>(f_src), static_cast<float *>(f_dst), sflags,
                           d);
    else if (pop_dtype == WAAM_POP_F16 || pop_dtype == WAAM_POP_BF16)
        hipLaunchKernelGGL(k_stream_pending<uint16_t>, grid3(d), block3(), 0, as_stream(stream),
                           static_cast<const uint16_t *>(f_src), static_cast<uint16_t *>(f_dst),
                           sflags, d);
    else
        return set_error(WAAM_ERR_INVALID, "waam_stream_pending: unknown population form %d",
                         pop_dtype);
    return check_launch("waam_stream_pending");
}

}  // extern "C"

namespace waam {

/* q-major population arrays: element e belongs to direction e / n_cells */
template <int PD>
__global__ void k_pack_pops(const float *__restrict__ src, uint16_t *__restrict__ dst, size_t n,
                            size_t n_cells)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        PopStore<PD>::store(dst, e, (int)(e / n_cells), src[e]);
}

template <int PD>
__global__ void k_unpack_pops(const uint16_t *__restrict__ src, float *__restrict__ dst, size_t n,
                              size_t n_cells)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        dst[e] = PopStore<PD>::load(src, e, (int)(e / n_cells));
}

/* ---- the droplet step (waam_step_fused_droplet) ---------------------------------------- */

/* arc heat and the ceiling applied to H in place on every non-GAS cell: the eager form of the
 * arc kernel 2 applies on the fly (reference coupled_step.py:103-108, before the droplet
 * block); the step's arc_k, the column scan of kernel 1 (tot[2]), is kept in tot[4] for the
 * rest of the step, whose column the deposition may raise */
__global__ void __launch_bounds__(256) k_droplet_arc(waam_step_fields f, StepConsts s)
{
    const Dims &d = s.d;
    if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0 &&
        threadIdx.y == 0) {
        float *tot = reinterpret_cast<float *>((char *)f.scratch + kScratchTelemetry);
        tot[4] = tot[2];
    }
    int i, j, k;
    if (!d.cell(i, j, k)) return;
    const HView hv = make_hview(f, s);
    const size_t c = d.idx(i, j, k);
    if (f.flags_in[c] == GAS) return;
    f.H_in[c] = hv.at(c, i, j, k);
}

/* the droplet's placed volume into the run's ledger, in binary64 and droplet order, as the
 * reference's host sums it (coupled_step.py:155-167) */
__global__ void k_droplet_ledger(const float *__restrict__ result, double drop_vol,
                                 double *__restrict__ ledger)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double placed = (double)result[0];
    ledger[0] += placed;
    if (placed < drop_vol * 0.98) ledger[1] += 1.0;
}

__global__ void k_zero3(float *__restrict__ a, float *__restrict__ b, float *__restrict__ c,
                        size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        a[e] = 0.0f;
        b[e] = 0.0f;
        c[e] = 0.0f;
    }
}

static int check_droplet(const char *who, const waam_step_fields &f, const waam_droplet &dp)
{
    WAAM_RETURN_IF(check_ptrs(who, {dp.f32_src, dp.f32_dst, dp.vol_acc, dp.real_acc, dp.result,
                                    dp.ledger}));
    if (f.pop_dtype == WAAM_POP_F32 && (dp.f32_src != f.f_src || dp.f32_dst != f.f_dst))
        return set_error(WAAM_ERR_INVALID, "%s: fp32 storage: f32_src/f32_dst must be f_src/f_dst",
                         who);
    if (f.pop_dtype != WAAM_POP_F32 &&
        (dp.f32_src == f.f_src || dp.f32_dst == f.f_dst || dp.f32_src == dp.f32_dst))
        return set_error(WAAM_ERR_INVALID, "%s: the fp32 pair aliases the 16-bit pair", who);
    if (dp.max_tracers > 0 && (!dp.tracer_pos || !dp.tracer_active || !dp.tracer_head))
        return set_error(WAAM_ERR_INVALID, "%s: tracer buffers missing", who);
    if (!(dp.drop_r >= 0.0f) || !(dp.drop_vol >= 0.0) || !(dp.footprint_r >= 0.0))
        return set_error(WAAM_ERR_INVALID, "%s: negative droplet size", who);
    return 0;
}

static void pops_convert(bool pack, void *p16, float *p32, size_t n_cells, int pd, hipStream_t st)
{
    const size_t n = n_cells * Q;
    unsigned g = grid_for(n);
    if (g > 65536u) g = 65536u;
    if (pack) {
        uint16_t *o = static_cast<uint16_t *>(p16);
        if (pd == WAAM_POP_F16)
            hipLaunchKernelGGL(k_pack_pops<WAAM_POP_F16>, dim3(g), dim3(kBlock), 0, st, p32, o, n,
                               n_cells);
        else
            hipLaunchKernelGGL(k_pack_pops<WAAM_POP_BF16>, dim3(g), dim3(kBlock), 0, st, p32, o, n,
                               n_cells);
    } else {
        const uint16_t *i = static_cast<const uint16_t *>(p16);
        if (pd == WAAM_POP_F16)
            hipLaunchKernelGGL(k_unpack_pops<WAAM_POP_F16>, dim3(g), dim3(kBlock), 0, st, i, p32,
                               n, n_cells);
        else
            hipLaunchKernelGGL(k_unpack_pops<WAAM_POP_BF16>, dim3(g), dim3(kBlock), 0, st, i, p32,
                               n, n_cells);
    }
}

/* One droplet step: the droplet block of the reference with arc_k, the placed volume and the
 * retry loop on the device, then the fused step with the arc already applied. The launch order
 * is the reference's order of these operations (solvers/coupled_step.py:97-166 and the fused
 * engine's eager-arc step). */
static void launch_droplet_step(const StepConsts &s0, const waam_step_fields &f,
                                const waam_droplet &dp, hipStream_t st, bool capture)
{
    const Dims &d = s0.d;
    const size_t N = d.n();
    const float *arc_k = reinterpret_cast<const float *>((char *)f.scratch + kScratchTelemetry) + 2;
    const bool f32 = f.pop_dtype == WAAM_POP_F32;
    /* the previous step's deferred stream */
    if (s0.p.pending_stream) {
        if (f32)
            hipLaunchKernelGGL(k_stream_pending<float>, grid3(d), block3(), 0, st,
                               static_cast<const float *>(f.f_src), static_cast<float *>(f.f_dst),
                               (const uint8_t *)f.sflags_prev, d);
        else
            hipLaunchKernelGGL(k_stream_pending<uint16_t>, grid3(d), block3(), 0, st,
                               static_cast<const uint16_t *>(f.f_src),
                               static_cast<uint16_t *>(f.f_dst), (const uint8_t *)f.sflags_prev,
                               d);
    }
    /* the deposition and the momentum impact write fp32 populations */
    if (!f32) {
        pops_convert(false, f.f_src, dp.f32_src, N, f.pop_dtype, st);
        pops_convert(false, f.f_dst, dp.f32_dst, N, f.pop_dtype, st);
    }
    /* kernel 1's column scan and arc weight sums, then arc heat and the ceiling on H */
    StepConsts sa = s0;
    sa.p.arc_applied = 0;
    sa.p.use_arc_k_override = 0;
    sa.p.quiet_skip = 0;
    sa.p.thermal_list_ready = 0;
    hipLaunchKernelGGL(k_step_begin, dim3(kReduceBlocks), dim3(kBlock), 0, st, f, sa, 1);
    hipLaunchKernelGGL(k_step_arc_total, dim3(1), dim3(kBlock), 0, st, f, sa);
    hipLaunchKernelGGL(k_droplet_arc, grid3(d), block3(), 0, st, f, sa);
    /* forces, cleared before the droplet's impact pressure adds to them */
    {
        unsigned g = grid_for(N);
        if (g > 65536u) g = 65536u;
        hipLaunchKernelGGL(k_zero3, dim3(g), dim3(kBlock), 0, st, f.Fx, f.Fy, f.Fz, N);
    }
    const float *arc_k4 = arc_k + 2;  // tot[4]: the scan before the deposition
    if (dp.max_tracers > 0 && dp.spawn > 0)
        (void)launch_tracers_dev(dp.tracer_pos, dp.tracer_active, dp.tracer_head, dp.max_tracers,
                                 dp.tracer_x, dp.tracer_y, arc_k4, dp.dx, dp.sigma_m, dp.spawn,
                                 dp.seed, st);
    (void)launch_deposit_dev(dp.f32_src, f.flags_in, f.fl_in, f.phi, f.H_in, f.T_in, f.rho,
                             dp.vol_acc, dp.real_acc, f.scratch, dp.result, dp.arc_i, dp.arc_j,
                             arc_k4, dp.footprint_r, dp.drop_r, dp.drop_vol, dp.T_drop, dp.cp_rho,
                             dp.L_rho, dp.rho0, dp.cell_vol, d, st);
    hipLaunchKernelGGL(k_droplet_ledger, dim3(1), dim3(64), 0, st, (const float *)dp.result,
                       dp.drop_vol, dp.ledger);
    if (dp.momentum)
        (void)launch_momentum_impact_dev(dp.f32_src, f.rho, f.ux, f.uy, f.uz, f.flags_in, f.fl_in,
                                         dp.arc_i, dp.arc_j, arc_k4, dp.drop_r, dp.vx, dp.vy,
                                         dp.vz, d, st);
    if (dp.pressure)
        (void)launch_droplet_pressure_dev(f.Fz, f.flags_in, f.phi, f.fl_in, dp.arc_i, dp.arc_j,
                                          arc_k4, dp.drop_r, dp.impact_pa, dp.dt, dp.dx_f,
                                          dp.rho_ref, d, st);
    if (!f32) {
        pops_convert(true, f.f_src, dp.f32_src, N, f.pop_dtype, st);
        pops_convert(true, f.f_dst, dp.f32_dst, N, f.pop_dtype, st);
    }
    /* the step, with the arc applied, the scan kept in tot[4] and the impact's forces */
    StepConsts s = s0;
    s.p.arc_applied = 1;
    s.p.use_arc_k_override = 2;
    s.p.forces_from_fields = 1;
    s.p.pending_stream = 0;
    s.p.thermal_list_ready = 0;
    s.p.force_reset_all = 0;
    s.p.quiet_skip = 0;
    launch_step(s, f, st, capture);
}

}  // namespace waam

extern "C" {

int waam_step_fused_droplet(const waam_step_params *pp, const waam_step_fields *ff,
                            const waam_droplet *dd, void *stream)
{
    if (!pp || !ff || !dd)
        return set_error(WAAM_ERR_INVALID, "waam_step_fused_droplet: NULL argument");
    const waam_step_params &p = *pp;
    const waam_step_fields &f = *ff;
    WAAM_RETURN_IF(check_fields("waam_step_fused_droplet", f));
    WAAM_RETURN_IF(check_tables("waam_step_fused_droplet", p, f));
    WAAM_RETURN_IF(check_droplet("waam_step_fused_droplet", f, *dd));
    if (p.col_i < 0 || p.col_i >= f.nx || p.col_j < 0 || p.col_j >= f.ny)
        return set_error(WAAM_ERR_INVALID, "waam_step_fused_droplet: arc column outside the grid");
    if (!p.is_welding || p.arc_kind > 1)
        return set_error(WAAM_ERR_INVALID,
                         "waam_step_fused_droplet: needs a welding step with the Gaussian or "
                         "Goldak source");
    Dims d{f.nx, f.ny, f.nz, f.layout};
    StepConsts s = make_step_consts(p, d);
    launch_droplet_step(s, f, *dd, as_stream(stream), false);
    return check_launch("waam_step_fused_droplet");
}

}  // extern "C"

extern "C" {

int waam_pack_populations(const float *src, void *dst, int64_t n_cells, int pop_dtype,
                          void *stream)
{
    WAAM_RETURN_IF(check_ptrs("waam_pack_populations", {src, dst}));
    if (n_cells <= 0)
        return set_error(WAAM_ERR_INVALID, "waam_pack_populations: n_cells %lld",
                         (long long)n_cells);
    size_t n = (size_t)n_cells * Q;
    unsigned g = grid_for(n);
    if (g > 65536u) g = 65536u;
    uint16_t *o = static_cast<uint16_t *>(dst);
    if (pop_dtype == WAAM_POP_F16)
        hipLaunchKernelGGL(k_pack_pops<WAAM_POP_F16>, dim3(g), dim3(kBlock), 0, as_stream(stream),
                           src, o, n, (size_t)n_cells);
    else if (pop_dtype == WAAM_POP_BF16)
        hipLaunchKernelGGL(k_pack_pops<WAAM_POP_BF16>, dim3(g), dim3(kBlock), 0, as_stream(stream),
                           src, o, n, (size_t)n_cells);
    else
        return set_error(WAAM_ERR_INVALID, "waam_pack_populations: pop_dtype %d is not a 16-bit "
                         "form", pop_dtype);
    return check_launch("waam_pack_populations");
}

int waam_unpack_populations(const void *src, float *dst, int64_t n_cells, int pop_dtype,
                            void *stream)
{
    WAAM_RETURN_IF(check_ptrs("waam_unpack_populations", {src, dst}));
    if (n_cells <= 0)
        return set_error(WAAM_ERR_INVALID, "waam_unpack_populations: n_cells %lld",
                         (long long)n_cells);
    size_t n = (size_t)n_cells * Q;
    unsigned g = grid_for(n);
    if (g > 65536u) g = 65536u;
    const uint16_t *i = static_cast<const uint16_t *>(src);
    if (pop_dtype == WAAM_POP_F16)
        hipLaunchKernelGGL(k_unpack_pops<WAAM_POP_F16>, dim3(g), dim3(kBlock), 0,
                           as_stream(stream), i, dst, n, (size_t)n_cells);
    else if (pop_dtype == WAAM_POP_BF16)
        hipLaunchKernelGGL(k_unpack_pops<WAAM_POP_BF16>, dim3(g), dim3(kBlock), 0,
                           as_stream(stream), i, dst, n, (size_t)n_cells);
    else
        return set_error(WAAM_ERR_INVALID, "waam_unpack_populations: pop_dtype %d is not a "
                         "16-bit form", pop_dtype);
    return check_launch("waam_unpack_populations");
}

int waam_step_droplet_arc_k(const void *scratch, float *host_out)
{
    if (!scratch || !host_out)
        return set_error(WAAM_ERR_INVALID, "waam_step_droplet_arc_k: NULL");
    hipError_t e = hipMemcpy(host_out, (const char *)scratch + kScratchTelemetry + 4 * sizeof(float),
                             sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess)
        return set_error(WAAM_ERR_HIP, "waam_step_droplet_arc_k: %s", hipGetErrorString(e));
    return 0;
}

int waam_step_arc_readback(const void *scratch, float *host_out3)
{
    if (!scratch || !host_out3) return set_error(WAAM_ERR_INVALID, "waam_step_arc_readback: NULL");
    hipError_t e = hipMemcpy(host_out3, (const char *)scratch + kScratchTelemetry, 3 * sizeof(float),
                             hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_error(WAAM_ERR_HIP, "waam_step_arc_readback: %s", hipGetErrorString(e));
    return WAAM_OK;
}

int waam_timing_begin(int which, int max_launches)
{
    return waam_timing_begin_every(which, max_launches, 1);
}

int waam_timing_begin_every(int which, int max_launches, int every)
{
    KernelTimer &t = g_timer;
    if (every <= 0)
        return set_error(WAAM_ERR_INVALID, "waam_timing_begin_every: every %d out of range", every);
    if (t.armed) return set_error(WAAM_ERR_INVALID, "waam_timing_begin: timing already armed");
    if (which < 0 || which >= WAAM_K_COUNT)
        return set_error(WAAM_ERR_INVALID, "waam_timing_begin: kernel id %d out of range", which);
    if (max_launches <= 0 || max_launches > (1 << 20))
        return set_error(WAAM_ERR_INVALID, "waam_timing_begin: max_launches %d out of range",
                         max_launches);
    t.ev.assign((size_t)2 * max_launches, nullptr);
    for (auto &e : t.ev) {
        hipError_t err = hipEventCreate(&e);
        if (err != hipSuccess) {
            for (auto &x : t.ev)
                if (x) (void)hipEventDestroy(x);
            t.ev.clear();
            return set_error(WAAM_ERR_HIP, "waam_timing_begin: %s", hipGetErrorString(err));
        }
    }
    t.which = which;
    t.used = 0;
    t.every = every;
    t.calls = 0;
    t.armed = true;
    return WAAM_OK;
}

int waam_timing_end(double *total_ms, int64_t *launches)
{
    KernelTimer &t = g_timer;
    if (!t.armed) return set_error(WAAM_ERR_INVALID, "waam_timing_end: timing not armed");
    double sum = 0.0;
    int rc = WAAM_OK;
    for (int i = 0; i < t.used; ++i) {
        float ms = 0.0f;
        hipError_t e = hipEventSynchronize(t.ev[2 * i + 1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t.ev[2 * i], t.ev[2 * i + 1]);
        if (e != hipSuccess) {
            rc = set_error(WAAM_ERR_HIP, "waam_timing_end: %s", hipGetErrorString(e));
            break;
        }
        sum += ms;
    }
    for (auto &e : t.ev) (void)hipEventDestroy(e);
    t.ev.clear();
    t.armed = false;
    if (total_ms) *total_ms = sum;
    if (launches) *launches = t.used;
    return rc;
}

}  // extern "C"
