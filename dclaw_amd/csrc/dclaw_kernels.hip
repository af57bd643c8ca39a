// dclaw_kernels.hip -- gfx950 kernels and the C-ABI of the D-Claw step.
//
// Kernels:
//   k_bc_fill        ghost-cell fill (GeoClaw bc2amr extrap / wall)
//   k_stepgrid       fused stepgrid for one patch: one 64-lane wavefront per
//                    (58-column strip, row segment); see dclaw_strip.h
//   k_step_batched   the same body over a work list of (patch, strip,
//                    segment) tiles of many patches in one launch
//   k_dt_update      dt_{n+1} = min(dt_n*cfl_desired/cfl_n, dt_max) per member
//   k_step2          GeoClaw step2 alone (fm, fp, gm, gp, cflgrid): the strip
//                    body in flux mode
//   k_b4step2_cells  b4step2's pointwise part (qfix, bdif >= 0, speed limit)
//   k_calc_taudir    calc_taudir (aux 4..6) on cells 0..mx+1 x 0..my+1
//   k_src2           src2 on the interior cells
//   k_check4nans     first NaN of q (AMRClaw check4nans), k_fill zero fill
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>

#include "dclaw_hip.h"
#include "dclaw_adjoint.h"
#include "dclaw_amr.h"
#include "dclaw_strip.h"
#include "dclaw_tangent.h"

using namespace dclaw;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    return fail(DCLAW_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                   \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

// Lane policy of the strip body on a 64-lane wavefront.  prev/next are
// whole-wave DPP shifts by one lane (wave_shr:1 = 0x138, wave_shl:1 =
// 0x130); the end lane whose source is outside the wave keeps its own value.
struct GpuLane {
    __device__ __forceinline__ int id() const { return (int)threadIdx.x; }
    template <int CTRL>
    __device__ __forceinline__ static double dpp(double v)
    {
        const int lo = __double2loint(v), hi = __double2hiint(v);
        const int rlo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xf, 0xf, false);
        const int rhi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xf, 0xf, false);
        return __hiloint2double(rhi, rlo);
    }
    __device__ __forceinline__ double prev(double v) const { return dpp<0x138>(v); }
    __device__ __forceinline__ double next(double v) const { return dpp<0x130>(v); }
    // value and tangent shift together (tangent-linear step)
    __device__ __forceinline__ dclaw_tl::Dual prev(const dclaw_tl::Dual& v) const
    {
        return dclaw_tl::Dual(prev(v.v), prev(v.d));
    }
    __device__ __forceinline__ dclaw_tl::Dual next(const dclaw_tl::Dual& v) const
    {
        return dclaw_tl::Dual(next(v.v), next(v.d));
    }
    __device__ __forceinline__ bool any(bool p) const { return __any((int)p) != 0; }
};

__device__ __forceinline__ void wave_max_to(double v, double* dst)
{
    for (int off = 32; off >= 1; off >>= 1) v = dmax(v, __shfl_xor(v, off, 64));
    if (threadIdx.x == 0 && v > 0.0)
        atomicMax(reinterpret_cast<unsigned long long*>(dst),
                  (unsigned long long)__double_as_longlong(v));
}

// Performance settings of the single-patch launches (results do not depend
// on them); atomic, so a thread may change one while others launch.
std::atomic<int> g_seg_rows{32};   // rows per segment of the single-patch launch
std::atomic<int> g_dry_rows{48};   // rows per wavefront of the split step's dry pass
std::atomic<int> g_split_mode{1};  // split step with the full workspace: 1 wet-first order,
                                   // 2 dry pass, 3 wet-first short segments + tall dry groups

// The settings one launch uses, read once per call: from the library-wide
// values above, or from a dclaw_handle's own.
struct Tuning {
    int seg_rows, split_mode, dry_rows;
};
inline Tuning global_tuning() { return Tuning{g_seg_rows.load(), g_split_mode.load(), g_dry_rows.load()}; }

// Tiles (58-column strip x segment) of a patch at a segment height.
inline long tiles_at(int mx, int my, int seg_rows)
{
    return (long)((mx + STRIP_W - 1) / STRIP_W) * (long)((my + seg_rows - 1) / seg_rows);
}

// Compute units of the current device (cached per device).
int cu_count()
{
    static int cache[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cache[dev]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n < 1)
            n = 256;
        cache[dev] = n;
    }
    return cache[dev];
}

template <int MODE>
__global__ __launch_bounds__(64) void k_stepgrid(const double* __restrict__ q_in,
                                                 double* __restrict__ q_out,
                                                 const double* __restrict__ aux,
                                                 double* aux_out, int maux, int mx, int my,
                                                 double dx, double dy,
                                                 const double* __restrict__ dt_dev,
                                                 const dclaw_params* __restrict__ params,
                                                 double* cfl_dev, int seg_rows, int nstrips,
                                                 int nseg, int bx0, int bx1, int by0, int by1,
                                                 int* tile_flags,
                                                 const int* __restrict__ order,
                                                 const int* __restrict__ dry_age)
{
    // tile order: row-major, or the permutation `order` (wet tiles of the
    // previous step first, so the long tiles start first); the dispatcher
    // deals consecutive workgroups to the XCDs round-robin, which spreads a
    // cluster of wet tiles over the whole chip
    const int T = order ? order[blockIdx.x] : (int)blockIdx.x;
    // a tile whose neighbourhood was dry on the last three inputs already
    // holds its result in q_out (k_dry_age); its flag is 0 and stays 0
    if (dry_age && dry_age[T] >= 3) return;
    const int strip = T % nstrips, seg = T / nstrips;
    const dclaw_params P = *params;
    StripArgs A;
    A.bc[0] = bx0;
    A.bc[1] = bx1;
    A.bc[2] = by0;
    A.bc[3] = by1;
    A.q_in = q_in;
    A.q_out = q_out;
    A.aux = aux;
    A.aux_out = aux_out;
    A.maux = maux;
    A.mx = mx;
    A.my = my;
    A.dx = dx;
    A.dy = dy;
    A.dt = *dt_dev;
    A.i0 = 1 + strip * STRIP_W;
    A.jA = 1 + seg * seg_rows;
    A.jB = min(my, A.jA + seg_rows - 1);
    A.fm = A.fp = A.gm = A.gp = nullptr;
    A.tile_flag = tile_flags ? tile_flags + (seg * nstrips + strip) : nullptr;
    GpuLane L;
    const double cfl = strip_step_dispatch<MODE>(L, P, A);
    wave_max_to(cfl, cfl_dev);
}

template <int MODE>
__global__ __launch_bounds__(64) void k_step_batched(const dclaw_patch* __restrict__ patches,
                                                     const int32_t* __restrict__ tiles,
                                                     int seg_rows,
                                                     const double* __restrict__ dt_dev,
                                                     const dclaw_params* __restrict__ params,
                                                     double* cfl_dev, int bx0, int bx1, int by0,
                                                     int by1, int* tile_flags,
                                                     const int* __restrict__ order,
                                                     const int* __restrict__ dry_age)
{
    const int t = order ? order[blockIdx.x] : (int)blockIdx.x;
    // a tile whose neighbourhood was dry on its patch's last three inputs
    // already holds its result in q_out (k_bdry_age); its flag stays 0
    if (dry_age && dry_age[t] >= 3) return;
    const int pidx = tiles[4 * t + 0];
    const dclaw_patch pt = patches[pidx];
    const dclaw_params P = params[pt.member];
    StripArgs A;
    A.q_in = pt.q_in;
    A.q_out = pt.q_out;
    A.aux = pt.aux;
    A.aux_out = nullptr;
    A.bc[0] = bx0;
    A.bc[1] = bx1;
    A.bc[2] = by0;
    A.bc[3] = by1;
    A.maux = pt.maux;
    A.mx = pt.mx;
    A.my = pt.my;
    A.dx = pt.dx;
    A.dy = pt.dy;
    A.dt = dt_dev[pt.member];
    A.i0 = 1 + tiles[4 * t + 1] * STRIP_W;
    A.jA = 1 + tiles[4 * t + 2] * seg_rows;
    A.jB = min(pt.my, A.jA + seg_rows - 1);
    A.fm = A.fp = A.gm = A.gp = nullptr;
    A.tile_flag = tile_flags ? tile_flags + t : nullptr;
    GpuLane L;
    const double cfl = strip_step_dispatch<MODE>(L, P, A);
    wave_max_to(cfl, cfl_dev + pt.member);
}

// Tangent-linear fused step (one patch, MODE_STEP): q_out and cflgrid equal
// k_stepgrid<MODE_STEP>'s bit for bit, dq_out = J_q dq_in + J_P dP at the
// launch's dt.
__global__ __launch_bounds__(64) void k_stepgrid_tl(const double* __restrict__ q_in,
                                                    const double* __restrict__ dq_in,
                                                    double* __restrict__ q_out,
                                                    double* __restrict__ dq_out,
                                                    const double* __restrict__ aux, int maux,
                                                    int mx, int my, double dx, double dy,
                                                    const double* __restrict__ dt_dev,
                                                    const dclaw_params* __restrict__ params,
                                                    const dclaw_params* __restrict__ dparams,
                                                    double* cfl_dev, int seg_rows, int nstrips,
                                                    int bx0, int bx1, int by0, int by1)
{
    const int strip = (int)blockIdx.x % nstrips, seg = (int)blockIdx.x / nstrips;
    const dclaw_tl::Params P = dclaw_tl::make_params(*params, *dparams);
    dclaw_tl::StripArgs A;
    A.q_in.v = q_in;
    A.q_in.d = dq_in;
    A.q_out.v = q_out;
    A.q_out.d = dq_out;
    A.aux = aux;
    A.aux_out = nullptr;
    A.maux = maux;
    A.mx = mx;
    A.my = my;
    A.dx = dx;
    A.dy = dy;
    A.dt = *dt_dev;
    A.i0 = 1 + strip * STRIP_W;
    A.jA = 1 + seg * seg_rows;
    A.jB = min(my, A.jA + seg_rows - 1);
    A.bc[0] = bx0;
    A.bc[1] = bx1;
    A.bc[2] = by0;
    A.bc[3] = by1;
    A.fm = A.fp = A.gm = A.gp = nullptr;
    A.tile_flag = nullptr;
    GpuLane L;
    const dclaw_tl::Dual cfl = dclaw_tl::strip_step<MODE_STEP>(L, P, A);
    wave_max_to(cfl.v, cfl_dev);
}

// Tangent-linear fused step over a batched work list: tangents[p] holds the
// tangent arrays of patch p (its q_in / q_out fields), dparams[member] the
// member's direction.
__global__ __launch_bounds__(64) void k_step_batched_tl(const dclaw_patch* __restrict__ patches,
                                                        const dclaw_patch* __restrict__ tangents,
                                                        const int32_t* __restrict__ tiles,
                                                        int seg_rows,
                                                        const double* __restrict__ dt_dev,
                                                        const dclaw_params* __restrict__ params,
                                                        const dclaw_params* __restrict__ dparams,
                                                        double* cfl_dev, int bx0, int bx1,
                                                        int by0, int by1)
{
    const int t = (int)blockIdx.x;
    const int pidx = tiles[4 * t + 0];
    const dclaw_patch pt = patches[pidx];
    const dclaw_patch tp = tangents[pidx];
    const dclaw_tl::Params P = dclaw_tl::make_params(params[pt.member], dparams[pt.member]);
    dclaw_tl::StripArgs A;
    A.q_in.v = pt.q_in;
    A.q_in.d = tp.q_in;
    A.q_out.v = pt.q_out;
    A.q_out.d = tp.q_out;
    A.aux = pt.aux;
    A.aux_out = nullptr;
    A.bc[0] = bx0;
    A.bc[1] = bx1;
    A.bc[2] = by0;
    A.bc[3] = by1;
    A.maux = pt.maux;
    A.mx = pt.mx;
    A.my = pt.my;
    A.dx = pt.dx;
    A.dy = pt.dy;
    A.dt = dt_dev[pt.member];
    A.i0 = 1 + tiles[4 * t + 1] * STRIP_W;
    A.jA = 1 + tiles[4 * t + 2] * seg_rows;
    A.jB = min(pt.my, A.jA + seg_rows - 1);
    A.fm = A.fp = A.gm = A.gp = nullptr;
    A.tile_flag = nullptr;
    GpuLane L;
    const dclaw_tl::Dual cfl = dclaw_tl::strip_step<MODE_STEP>(L, P, A);
    wave_max_to(cfl.v, cfl_dev + pt.member);
}

// src2 over the interior cells of the tiles the no-src2 step flagged (one
// patch; tile t = segment * nstrips + strip).  One wavefront per task, a task
// being a 64-cell chunk of a tile, so the wet tiles spread over the whole chip;
// grid-stride over the tasks.
__global__ __launch_bounds__(64, 4) void k_src2_tiles(double* q, const double* __restrict__ aux,
                                                   int maux, int mx, int my, int seg_rows,
                                                   int nstrips, int ntiles,
                                                   const int* __restrict__ tile_flags,
                                                   const double* __restrict__ dt_dev,
                                                   const dclaw_params* __restrict__ params)
{
    const dclaw_params P = *params;
    const double dt = *dt_dev;
    const int nchunk = (STRIP_W * seg_rows + 63) / 64;
    const long ntask = (long)ntiles * nchunk;
    for (long task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int t = (int)(task / nchunk);
        if (!tile_flags[t]) continue;
        const int i0 = 1 + (t % nstrips) * STRIP_W;
        const int jA = 1 + (t / nstrips) * seg_rows;
        const int jB = min(my, jA + seg_rows - 1);
        const int k = (int)(task % nchunk) * 64 + (int)threadIdx.x;
        if (k < STRIP_W * (jB - jA + 1)) src2_tile_cell(P, dt, q, aux, maux, mx, i0, jA, jB, k);
    }
}

// Dry pass of the split step (one patch): one wavefront per group of `group`
// consecutive segments of a strip.  It writes the dry state of every cell the
// group owns and src2 flag 0 for its dry segments, and appends the others
// (tile index segment * nstrips + strip) to the wet list, counters[0] being
// the list length (zero before the launch).
__global__ __launch_bounds__(64) void k_classify(const double* __restrict__ q_in,
                                                 double* __restrict__ q_out, int mx, int my,
                                                 const dclaw_params* __restrict__ params,
                                                 int seg_rows, int group, int nseg,
                                                 int nstrips, int bx0, int bx1,
                                                 int by0, int by1, int* tile_flags,
                                                 int* wet_list, int* counters)
{
    const int strip = (int)blockIdx.x % nstrips, gi = (int)blockIdx.x / nstrips;
    StripArgs A;
    A.bc[0] = bx0;
    A.bc[1] = bx1;
    A.bc[2] = by0;
    A.bc[3] = by1;
    A.q_in = q_in;
    A.q_out = q_out;
    A.aux = nullptr;
    A.aux_out = nullptr;
    A.maux = 1;
    A.mx = mx;
    A.my = my;
    A.dx = A.dy = A.dt = 0.0;
    A.i0 = 1 + strip * STRIP_W;
    const int g0 = gi * group;
    A.jA = 1 + g0 * seg_rows;
    A.jB = min(my, A.jA + group * seg_rows - 1);
    A.fm = A.fp = A.gm = A.gp = nullptr;
    A.tile_flag = nullptr;
    GpuLane L;
    const unsigned wet = strip_classify(L, params->dry_tolerance, A, seg_rows);
    if (threadIdx.x == 0) {
        const int ng = min(group, nseg - g0);
        const int nw = __popc(wet);
        int at = nw ? atomicAdd(counters, nw) : 0;
        for (int s = 0; s < ng; s++) {
            const int t = (g0 + s) * nstrips + strip;
            if ((wet >> s) & 1u) wet_list[at++] = t;
            else tile_flags[t] = 0;
        }
    }
}

// The strip body over the wet list: persistent wavefronts take list entries
// through the work cursor counters[1] (zero before the launch) until the
// list (counters[0] entries) is exhausted.
template <int MODE>
__global__ __launch_bounds__(64) void k_stepgrid_list(const double* __restrict__ q_in,
                                                      double* __restrict__ q_out,
                                                      const double* __restrict__ aux, int maux,
                                                      int mx, int my, double dx, double dy,
                                                      const double* __restrict__ dt_dev,
                                                      const dclaw_params* __restrict__ params,
                                                      double* cfl_dev, int seg_rows, int nstrips,
                                                      int bx0, int bx1, int by0, int by1,
                                                      int* tile_flags,
                                                      const int* __restrict__ wet_list,
                                                      int* counters)
{
    const dclaw_params P = *params;
    StripArgs A;
    A.bc[0] = bx0;
    A.bc[1] = bx1;
    A.bc[2] = by0;
    A.bc[3] = by1;
    A.q_in = q_in;
    A.q_out = q_out;
    A.aux = aux;
    A.aux_out = nullptr;
    A.maux = maux;
    A.mx = mx;
    A.my = my;
    A.dx = dx;
    A.dy = dy;
    A.dt = *dt_dev;
    A.fm = A.fp = A.gm = A.gp = nullptr;
    A.known_wet = true;
    const int n = counters[0];
    GpuLane L;
    double cfl = 0.0;
    for (;;) {
        int w = 0;
        if (threadIdx.x == 0) w = atomicAdd(counters + 1, 1);
        w = __builtin_amdgcn_readfirstlane(w);
        if (w >= n) break;
        const int t = wet_list[w];
        A.i0 = 1 + (t % nstrips) * STRIP_W;
        A.jA = 1 + (t / nstrips) * seg_rows;
        A.jB = min(my, A.jA + seg_rows - 1);
        A.tile_flag = tile_flags + t;
        cfl = dmax(cfl, strip_step_dispatch<MODE>(L, P, A));
    }
    wave_max_to(cfl, cfl_dev);
}

// src2 over the interior cells of the wet-list tiles the strip body flagged;
// tasks are 64-cell chunks of those tiles, grid-strided.
__global__ __launch_bounds__(64) void k_src2_list(double* q, const double* __restrict__ aux,
                                                  int maux, int mx, int my, int seg_rows,
                                                  int nstrips, const int* __restrict__ tile_flags,
                                                  const int* __restrict__ wet_list,
                                                  const int* __restrict__ counters,
                                                  const double* __restrict__ dt_dev,
                                                  const dclaw_params* __restrict__ params)
{
    const int nchunk = (STRIP_W * seg_rows + 63) / 64;
    const long ntask = (long)counters[0] * nchunk;
    if ((long)blockIdx.x >= ntask) return;
    const dclaw_params P = *params;
    const double dt = *dt_dev;
    for (long task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int t = wet_list[task / nchunk];
        if (!tile_flags[t]) continue;
        const int i0 = 1 + (t % nstrips) * STRIP_W;
        const int jA = 1 + (t / nstrips) * seg_rows;
        const int jB = min(my, jA + seg_rows - 1);
        const int k = (int)(task % nchunk) * 64 + (int)threadIdx.x;
        if (k < STRIP_W * (jB - jA + 1)) src2_tile_cell(P, dt, q, aux, maux, mx, i0, jA, jB, k);
    }
}

// Dry-tile skipping of the ordered split step (a "resident" workspace of
// split_resident_ints(ntiles) ints: flags, order, counters, then SKIP_META
// ints of bookkeeping and one age per tile).  The strip kernel's tile flag
// marks a tile whose output has a cell wet before src2 (src2 only dries
// cells), so after a step the flags cover every wet cell of the new state.
// If no tile within the strip kernel's read halo of tile T (one strip, and
// ceil(3 / segment rows) segments, each way) is flagged, every cell the
// kernel reads for T is dry and it writes the dry map of its input, whose
// repeated application changes nothing.  Three such inputs in a row (D_n,
// D_{n-1}, D_{n-2}), with q_in and q_out swapping roles between calls,
// leave q_out's tile (step n-1's input) equal to the result, so the tile is
// not read or written.  meta: [0] segment rows, [1] strips, [2] segments,
// [3..4] the previous call's q_out, [5..6] its q_in, [7] how many
// consecutive calls (capped at 4) continued one such sequence -- the flags
// describe this call's input only when it is at least 1.
constexpr int SKIP_META = 8;
inline long split_resident_ints(long ntiles) { return 4 * ntiles + 4 + SKIP_META; }

__global__ void k_skip_meta(int* meta, const double* q_in, const double* q_out, int seg_rows,
                            int nstrips, int nseg)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const unsigned long long in = (unsigned long long)q_in, out = (unsigned long long)q_out;
    const unsigned long long last_out =
        (unsigned long long)(unsigned)meta[3] | ((unsigned long long)(unsigned)meta[4] << 32);
    const unsigned long long last_in =
        (unsigned long long)(unsigned)meta[5] | ((unsigned long long)(unsigned)meta[6] << 32);
    const bool cont = meta[0] == seg_rows && meta[1] == nstrips && meta[2] == nseg &&
                      in == last_out && out == last_in && meta[7] >= 0;
    meta[7] = cont ? min(meta[7] + 1, 4) : 0;
    meta[0] = seg_rows;
    meta[1] = nstrips;
    meta[2] = nseg;
    meta[3] = (int)(unsigned)(out & 0xffffffffull);
    meta[4] = (int)(unsigned)(out >> 32);
    meta[5] = (int)(unsigned)(in & 0xffffffffull);
    meta[6] = (int)(unsigned)(in >> 32);
}

// Whether the part of neighbour tile (dx, dy) (strips, segments; (0, 0)
// the tile itself) that a tile's march reads holds no cell flagged wet, from
// the neighbour's flag: with segments of at least STRIP_HALO rows a march
// reads 3 columns of each side neighbour, 3 rows of the segments below and
// above, and their corners -- the neighbour's bands facing the tile.
__device__ __forceinline__ bool halo_band_dry(int f, int dx, int dy)
{
    if (dx == 0 && dy == 0) return f == 0;
    const int bx = dx < 0 ? TF_EAST : TF_WEST;      // the columns facing the tile
    const int by = dy < 0 ? TF_NORTH : TF_SOUTH;    // the rows facing the tile
    if (dy == 0) return (f & bx) == 0;
    if (dx == 0) return (f & by) == 0;
    return (f & bx) == 0 || (f & by) == 0;          // the corner lies in both bands
}

// Age of each tile: consecutive inputs with no flagged cell within the read
// halo (0 when the flags do not describe this call's input).
__global__ void k_dry_age(const int* __restrict__ flags, int* age, const int* __restrict__ meta,
                          int nstrips, int nseg, int ry)
{
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= nstrips * nseg) return;
    if (meta[7] == 0) {
        age[t] = 0;
        return;
    }
    const int strip = t % nstrips, seg = t / nstrips;
    bool dry = true;
    if (ry == 1) {
        for (int s = max(0, seg - 1); s <= min(nseg - 1, seg + 1); s++)
            for (int x = max(0, strip - 1); x <= min(nstrips - 1, strip + 1); x++)
                dry = dry && halo_band_dry(flags[s * nstrips + x], x - strip, s - seg);
    } else {
        for (int s = max(0, seg - ry); s <= min(nseg - 1, seg + ry); s++)
            for (int x = max(0, strip - 1); x <= min(nstrips - 1, strip + 1); x++)
                dry = dry && flags[s * nstrips + x] == 0;
    }
    age[t] = dry ? min(age[t] + 1, 3) : 0;
}

// Permutation of the n tiles with the flagged ones first (flags from the
// previous step's strip kernel): counters[0] counts flagged tiles from the
// front, counters[1] the others from the back (both zero before the launch).
__global__ void k_order_tiles(const int* __restrict__ flags, int n, int* order, int* counters)
{
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const bool in = t < n;
    const bool wet = in && flags[t] != 0;
    const unsigned long long bw = __ballot(wet), bd = __ballot(in && !wet);
    const int lane = (int)(threadIdx.x & 63u);
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int base_w = 0, base_d = 0;
    if (lane == 0) {
        if (bw) base_w = atomicAdd(counters, __popcll(bw));
        if (bd) base_d = atomicAdd(counters + 1, __popcll(bd));
    }
    base_w = __shfl(base_w, 0, 64);
    base_d = __shfl(base_d, 0, 64);
    if (wet) order[base_w + __popcll(bw & below)] = t;
    else if (in) order[n - 1 - (base_d + __popcll(bd & below))] = t;
}

// Dry-tile skipping of the batched split step (n_ws >= batched_resident_ints
// in the workspace: flags, order, counters, then BSKIP_GMETA ints of the
// work list's shape, BSKIP_PMETA ints per patch and one age per tile).  Per
// patch the rule of the single-patch step above: the tile (patch, strip,
// segment) of a work list from dclaw_worklist_build sits at the patch's
// first tile + strip * segments + segment, and a tile whose neighbours
// within the strip kernel's read halo (one strip, ceil(3 / segment rows)
// segments) were unflagged on the last three inputs of that patch -- with
// its q_in and q_out swapping roles between calls -- holds its result in
// q_out already.  A neighbour is used only where the list has it at that
// place (any other list only skips nothing).  gmeta: [0] segment rows,
// [1] tiles, [2] patches, [3] 1 once written; pmeta per patch: [0..1] the
// previous call's q_out, [2..3] its q_in, [4] consecutive calls (capped at
// 4) that continued one such sequence.
constexpr int BSKIP_GMETA = 4;
constexpr int BSKIP_PMETA = 5;
inline long batched_resident_ints(long ntiles, long npatch)
{
    return 3 * ntiles + 4 + BSKIP_GMETA + BSKIP_PMETA * npatch;
}

__global__ __launch_bounds__(256) void k_bskip_meta(const dclaw_patch* __restrict__ patches,
                                                    int npatch, int* gmeta, int* pmeta,
                                                    int seg_rows, int ntiles)
{
    __shared__ int same;
    if (threadIdx.x == 0)
        same = gmeta[3] == 1 && gmeta[0] == seg_rows && gmeta[1] == ntiles && gmeta[2] == npatch;
    __syncthreads();
    for (int p = (int)threadIdx.x; p < npatch; p += (int)blockDim.x) {
        int* m = pmeta + BSKIP_PMETA * p;
        const unsigned long long in = (unsigned long long)patches[p].q_in;
        const unsigned long long out = (unsigned long long)patches[p].q_out;
        const unsigned long long last_out =
            (unsigned long long)(unsigned)m[0] | ((unsigned long long)(unsigned)m[1] << 32);
        const unsigned long long last_in =
            (unsigned long long)(unsigned)m[2] | ((unsigned long long)(unsigned)m[3] << 32);
        const bool cont = same && in == last_out && out == last_in && m[4] >= 0;
        m[4] = cont ? min(m[4] + 1, 4) : 0;
        m[0] = (int)(unsigned)(out & 0xffffffffull);
        m[1] = (int)(unsigned)(out >> 32);
        m[2] = (int)(unsigned)(in & 0xffffffffull);
        m[3] = (int)(unsigned)(in >> 32);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        gmeta[0] = seg_rows;
        gmeta[1] = ntiles;
        gmeta[2] = npatch;
        gmeta[3] = 1;
    }
}

__global__ void k_bdry_age(const dclaw_patch* __restrict__ patches,
                           const int32_t* __restrict__ tiles, const int* __restrict__ flags,
                           int* age, const int* __restrict__ pmeta, int ntiles, int seg_rows,
                           int ry)
{
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= ntiles) return;
    const int p = tiles[4 * t + 0], s = tiles[4 * t + 1], g = tiles[4 * t + 2];
    if (pmeta[BSKIP_PMETA * p + 4] == 0) {
        age[t] = 0;
        return;
    }
    const int ns = (patches[p].mx + STRIP_W - 1) / STRIP_W;
    const int ng = (patches[p].my + seg_rows - 1) / seg_rows;
    const long base = (long)t - ((long)s * ng + g);
    bool dry = true;
    for (int sg = max(0, g - ry); sg <= min(ng - 1, g + ry); sg++)
        for (int x = max(0, s - 1); x <= min(ns - 1, s + 1); x++) {
            const long nb = base + (long)x * ng + sg;
            const bool placed = nb >= 0 && nb < ntiles && tiles[4 * nb + 0] == p &&
                                tiles[4 * nb + 1] == x && tiles[4 * nb + 2] == sg;
            dry = dry && placed &&
                  (ry == 1 ? halo_band_dry(flags[nb], x - s, sg - g) : flags[nb] == 0);
        }
    age[t] = dry ? min(age[t] + 1, 3) : 0;
}

// Work list of the hybrid split step: each group of HYB_GROUP consecutive
// segments of a strip goes to the wet list as its single segments (tile index
// segment * nstrips + strip) if the previous step flagged any of them, else
// to the dry list as one entry (group index gi * nstrips + strip).
// counters[0], counters[1]: the two lengths (zero before the launch).
constexpr int HYB_GROUP = 4;
__global__ void k_order_hybrid(const int* __restrict__ flags, int nstrips, int nseg,
                               int* wet_list, int* dry_list, int* counters)
{
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int ngi = (nseg + HYB_GROUP - 1) / HYB_GROUP;
    if (g >= nstrips * ngi) return;
    const int strip = g % nstrips, gi = g / nstrips;
    const int s0 = gi * HYB_GROUP, s1 = min(nseg, s0 + HYB_GROUP);
    bool wet = false;
    for (int s = s0; s < s1; s++) wet |= flags[s * nstrips + strip] != 0;
    if (wet) {
        const int base = atomicAdd(counters, s1 - s0);
        for (int s = s0; s < s1; s++) wet_list[base + s - s0] = s * nstrips + strip;
    } else {
        dry_list[atomicAdd(counters + 1, 1)] = g;
    }
}

// The no-src2 strip kernel over the hybrid work list: workgroups [0, nw) take
// the wet list's single segments (the long tiles start first, and short ones
// keep the longest tile's march short), [nw, nw + nd) the dry groups as one
// tall segment each (one pass over their rows and halo); the rest exit.  A
// group's src2 flag is copied to each of its segments (a segment without a wet
// cell flagged with its group only costs src2 a no-op pass).
__global__ __launch_bounds__(64) void k_stepgrid_hybrid(const double* __restrict__ q_in,
                                                        double* __restrict__ q_out,
                                                        const double* __restrict__ aux, int maux,
                                                        int mx, int my, double dx, double dy,
                                                        const double* __restrict__ dt_dev,
                                                        const dclaw_params* __restrict__ params,
                                                        double* cfl_dev, int seg_rows, int nstrips,
                                                        int nseg, int bx0, int bx1, int by0,
                                                        int by1, int* tile_flags,
                                                        const int* __restrict__ wet_list,
                                                        const int* __restrict__ dry_list,
                                                        const int* __restrict__ counters)
{
    const int nw = counters[0], nd = counters[1];
    const int b = (int)blockIdx.x;
    int strip, s0, nsub;
    if (b < nw) {
        const int t = wet_list[b];
        strip = t % nstrips;
        s0 = t / nstrips;
        nsub = 1;
    } else if (b < nw + nd) {
        const int g = dry_list[b - nw];
        strip = g % nstrips;
        s0 = (g / nstrips) * HYB_GROUP;
        nsub = min(nseg, s0 + HYB_GROUP) - s0;
    } else {
        return;
    }
    const dclaw_params P = *params;
    StripArgs A;
    A.bc[0] = bx0;
    A.bc[1] = bx1;
    A.bc[2] = by0;
    A.bc[3] = by1;
    A.q_in = q_in;
    A.q_out = q_out;
    A.aux = aux;
    A.aux_out = nullptr;
    A.maux = maux;
    A.mx = mx;
    A.my = my;
    A.dx = dx;
    A.dy = dy;
    A.dt = *dt_dev;
    A.i0 = 1 + strip * STRIP_W;
    A.jA = 1 + s0 * seg_rows;
    A.jB = min(my, (s0 + nsub) * seg_rows);
    A.fm = A.fp = A.gm = A.gp = nullptr;
    int* flag0 = tile_flags + (s0 * nstrips + strip);
    A.tile_flag = flag0;
    GpuLane L;
    const double cfl = strip_step_dispatch<MODE_STEP_NOSRC2>(L, P, A);
    wave_max_to(cfl, cfl_dev);
    if (nsub > 1 && threadIdx.x == 0) {
        // the group's bands are not its segments': every band of each
        // segment counts as wet
        const int v = *flag0 ? TF_ALL : 0;
        for (int k = 0; k < nsub; k++) flag0[k * nstrips] = v;
    }
}

__global__ void k_zero_ints(int* a, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) a[i] = 0;
}

// The same over a batched work list (each tile's patch, parameters and dt).
__global__ __launch_bounds__(64, 4) void k_src2_batched(const dclaw_patch* __restrict__ patches,
                                                     const int32_t* __restrict__ tiles,
                                                     int ntiles, int seg_rows,
                                                     const int* __restrict__ tile_flags,
                                                     const double* __restrict__ dt_dev,
                                                     const dclaw_params* __restrict__ params)
{
    const int nchunk = (STRIP_W * seg_rows + 63) / 64;
    const long ntask = (long)ntiles * nchunk;
    for (long task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int t = (int)(task / nchunk);
        if (!tile_flags[t]) continue;
        const dclaw_patch pt = patches[tiles[4 * t + 0]];
        const dclaw_params P = params[pt.member];
        const double dt = dt_dev[pt.member];
        const int i0 = 1 + tiles[4 * t + 1] * STRIP_W;
        const int jA = 1 + tiles[4 * t + 2] * seg_rows;
        const int jB = min(pt.my, jA + seg_rows - 1);
        const int k = (int)(task % nchunk) * 64 + (int)threadIdx.x;
        if (k < STRIP_W * (jB - jA + 1))
            src2_tile_cell(P, dt, pt.q_out, pt.aux, pt.maux, pt.mx, i0, jA, jB, k);
    }
}

__global__ __launch_bounds__(64) void k_step2(const double* __restrict__ q,
                                              const double* __restrict__ aux, int maux, int mx,
                                              int my, double dx, double dy,
                                              const double* __restrict__ dt_dev,
                                              const dclaw_params* __restrict__ params,
                                              double* fm, double* fp, double* gm, double* gp,
                                              double* cfl_dev, int seg_rows)
{
    const dclaw_params P = *params;
    StripArgs A;
    A.bc[0] = A.bc[1] = A.bc[2] = A.bc[3] = 0;
    A.q_in = q;
    A.q_out = nullptr;
    A.aux = aux;
    A.aux_out = nullptr;
    A.maux = maux;
    A.mx = mx;
    A.my = my;
    A.dx = dx;
    A.dy = dy;
    A.dt = *dt_dev;
    A.i0 = 1 + (int)blockIdx.x * STRIP_W;
    A.jA = 1 + (int)blockIdx.y * seg_rows;
    A.jB = min(my, A.jA + seg_rows - 1);
    A.fm = fm;
    A.fp = fp;
    A.gm = gm;
    A.gp = gp;
    A.tile_flag = nullptr;
    GpuLane L;
    const double cfl = strip_step_dispatch<MODE_FLUX>(L, P, A);
    wave_max_to(cfl, cfl_dev);
}

// b4step2.f90:56-100 on every cell (ghosts included), in place.
__global__ void k_b4step2_cells(double* q, int mx, int my, const dclaw_params* __restrict__ params)
{
    const dclaw_params P = *params;
    const long n = (long)(mx + 4) * (my + 4);
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        double* c = q + 7 * t;
        double h = c[0], hu = c[1], hv = c[2], hm = c[3], p = c[4], hchi = c[5];
        double u, v, m, chi, rho;
        qfix(P, h, hu, hv, hm, p, hchi, u, v, m, chi, rho, P.grav);
        const double bdif = dmax(c[6], 0.0);
        if (h > 0.0) {
            const double sp = sqrt((hu * hu + hv * hv)) / h;
            if (sp > P.speed_limit) {
                const double sratio = P.speed_limit / sp;
                hu = hu * sratio;
                hv = hv * sratio;
            }
        }
        c[0] = h; c[1] = hu; c[2] = hv; c[3] = hm; c[4] = p; c[5] = hchi; c[6] = bdif;
    }
}

// calc_taudir (digclaw_module.f90:500-719) on cells 0..mx+1 x 0..my+1: reads
// the b4step2 output and aux(1), writes aux(4..6).  The outer ghost ring is
// left as it is, as in the reference.
__global__ void k_calc_taudir(const double* __restrict__ q, double* aux, int maux, int mx,
                              int my, double dx, double dy,
                              const dclaw_params* __restrict__ params)
{
    const dclaw_params P = *params;
    const int nx = mx + 2;
    const long n = (long)nx * (my + 2);
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx), j = (int)(t / nx);
        auto cell = [&](int ii, int jj, double& b) {
            const double* c = q + qidx(0, ii, jj, mx);
            b = aux[aidx(ii, jj, mx, maux)] - c[6];
            Cell r = {c[0], c[1], c[2], c[3], c[4], c[5]};
            return r;
        };
        double bC, bL, bR, bB, bT;
        const Cell C = cell(i, j, bC), Lc = cell(i - 1, j, bL), Rc = cell(i + 1, j, bR);
        const Cell Bc = cell(i, j - 1, bB), Tc = cell(i, j + 1, bT);
        double h = C.h, hu = C.hu, hv = C.hv, hm = C.hm, p = C.p, hchi = C.hchi;
        double u, v, m, chi, rho;
        qfix(P, h, hu, hv, hm, p, hchi, u, v, m, chi, rho, P.grav);
        RhoTau sv;
        sv.rho = 0.0;
        sv.tau = 0.0;
        if (h > P.dry_tolerance) sv = rho_tau(P, h, u, v, m, p, chi);
        const Taudir td = calc_taudir_cell(P, dx, dy, C, bC, Lc, bL, Rc, bR, Bc, bB, Tc, bT, sv);
        const size_t o = aidx(i, j, mx, maux);
        aux[o + 3] = td.fsphi;
        aux[o + 4] = td.tx;
        aux[o + 5] = td.ty;
    }
}

// src2.f90:92-300 on the interior cells, in place.
__global__ void k_src2(double* q, const double* __restrict__ aux, int maux, int mx, int my,
                       const double* __restrict__ dt_dev,
                       const dclaw_params* __restrict__ params)
{
    const dclaw_params P = *params;
    const double dt = *dt_dev;
    const long n = (long)mx * my;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(t % mx), j = 1 + (int)(t / mx);
        double* c = q + qidx(0, i, j, mx);
        double h = c[0], hu = c[1], hv = c[2], hm = c[3], p = c[4], hchi = c[5], bdif = c[6];
        const double h_e = P.entrainment ? aux[aidx(i, j, mx, maux) + 6] : 0.0;
        src2_cell(P, dt, h, hu, hv, hm, p, hchi, bdif, h_e);
        c[0] = h; c[1] = hu; c[2] = hv; c[3] = hm; c[4] = p; c[5] = hchi; c[6] = bdif;
    }
}

// Smallest flat index (component fastest) of a NaN in q, min-reduced into
// *first (which the caller initialises to INT32_MAX).
// ---- curvature (src2.f90:133-156): gz runs over src2's sweep
// The reference sets gz once and adds each wet cell's curvature acceleration
// to it without a reset, so a cell's gravity includes every earlier cell's
// in loop order (j outer, i inner).  The step forms that running sum exactly
// in the reference's order: the accelerations of all cells in parallel, the
// sum in one workgroup (a cell that adds nothing adds +0.0, which leaves the
// sum bit-identical), then src2 on every cell in parallel with its own gz.
__global__ void k_curv_gacc(const double* __restrict__ q, const double* __restrict__ aux,
                            int maux, int mx, int my, double dx, double dy,
                            const dclaw_params* __restrict__ params, double* __restrict__ g)
{
    const dclaw_params P = *params;
    const long n = (long)mx * my;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(t % mx), j = 1 + (int)(t / mx);
        g[t] = curvature_gacc(P, q, aux, maux, mx, i, j, dx, dy);
    }
}

// g[t] <- grav + g[0] + ... + g[t], added one by one in index order (one
// workgroup; chunks of 1024 through LDS; a chunk of zeros is skipped, since
// adding +0.0 does not change the sum).
__global__ __launch_bounds__(64) void k_curv_prefix(double* g, long n,
                                                    const dclaw_params* __restrict__ params)
{
    __shared__ double buf[1024];
    double gz = params->grav;
    for (long base = 0; base < n; base += 1024) {
        const int cnt = (int)min(1024L, n - base);
        bool nz = false;
        for (int k = threadIdx.x; k < cnt; k += 64) {
            const double v = g[base + k];
            buf[k] = v;
            nz = nz || v != 0.0;
        }
        __syncthreads();
        const bool any = __ballot(nz) != 0ull;   // uniform over the wavefront
        if (any && threadIdx.x == 0) {
            for (int k = 0; k < cnt; k++) {
                gz = gz + buf[k];
                buf[k] = gz;
            }
        }
        __syncthreads();
        gz = __shfl(gz, 0);
        for (int k = threadIdx.x; k < cnt; k += 64) g[base + k] = any ? buf[k] : gz;
        __syncthreads();
    }
}

__global__ void k_src2_curv(double* q, const double* __restrict__ aux, int maux, int mx, int my,
                            const double* __restrict__ dt_dev,
                            const dclaw_params* __restrict__ params,
                            const double* __restrict__ gz)
{
    const dclaw_params P = *params;
    const double dt = *dt_dev;
    const long n = (long)mx * my;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(t % mx), j = 1 + (int)(t / mx);
        double* c = q + qidx(0, i, j, mx);
        double h = c[0], hu = c[1], hv = c[2], hm = c[3], p = c[4], hchi = c[5], bdif = c[6];
        src2_cell_gz(P, dt, h, hu, hv, hm, p, hchi, bdif, 0.0, gz[t]);
        c[0] = h; c[1] = hu; c[2] = hv; c[3] = hm; c[4] = p; c[5] = hchi; c[6] = bdif;
    }
}

__global__ void k_check4nans(const double* __restrict__ q, long n, int* first)
{
    int mine = 0x7fffffff;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x)
        if (q[t] != q[t] && t < mine) mine = (int)t;
    if (mine != 0x7fffffff) atomicMin(first, mine);
}

__global__ void k_fill(double* a, long n, double v)
{
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x)
        a[t] = v;
}

__global__ void k_fill_int(int* a, int v) { *a = v; }

int grid_for(long n)
{
    return (int)std::max<long>(1, std::min<long>((n + 255) / 256, 8192));
}

// Ghost fill: each ghost cell copies its source cell directly (extrap: nearest
// interior cell; wall: mirror cell with the normal momentum negated).  The
// direct index map equals bc2amr's x-then-y fill for these two rules.
__global__ void k_bc_fill(double* q, int mx, int my, int bxl, int bxh, int byl, int byh)
{
    const int nx = mx + 4, ny = my + 4;
    const long n = (long)nx * ny;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int ii = (int)(t % nx) - 1;   // Fortran i in [-1, mx+2]
        const int jj = (int)(t / nx) - 1;
        if (ii >= 1 && ii <= mx && jj >= 1 && jj <= my) continue;
        const int bc[4] = {bxl, bxh, byl, byh};
        int si, sj;
        double fu, fv;
        bc_source(bc, mx, my, ii, jj, si, sj, fu, fv);
        const size_t src = qidx(0, si, sj, mx), dst = qidx(0, ii, jj, mx);
        for (int m = 0; m < 7; m++) {
            double val = q[src + m];
            if (m == 1) val = fu * val;
            if (m == 2) val = fv * val;
            q[dst + m] = val;
        }
    }
}

__global__ void k_dt_update(double* dt, double* cfl, int n, double cfl_desired, double dt_max)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double c = cfl[k];
    double dtn;
    if (c > 0.0) dtn = dt[k] * cfl_desired / c;
    else dtn = dt_max;
    dt[k] = dmin(dtn, dt_max);
    cfl[k] = 0.0;
}

// dt rule with the output-time hit of tick.f:135-144: the member has just
// taken a step of dt[k] at time t[k]; t advances, the next dt is the CFL
// rule capped by dt_max, shortened to land on tout[k] when
// t + 1.001 dt >= tout[k] (hit[k] = 1).
__global__ void k_dt_update_out(double* dt, double* cfl, double* t, const double* tout, int* hit,
                                int n, double cfl_desired, double dt_max)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double tn = t[k] + dt[k];
    const double c = cfl[k];
    double dtn;
    if (c > 0.0) dtn = dt[k] * cfl_desired / c;
    else dtn = dt_max;
    dtn = dmin(dtn, dt_max);
    int h = 0;
    const double to = tout[k];
    if (tn < to && tn + 1.001 * dtn >= to) {
        dtn = to - tn;
        h = 1;
    }
    t[k] = tn;
    dt[k] = dtn;
    cfl[k] = 0.0;
    hit[k] = h;
}

// Gauges: q components and eta = h + (aux(1) - bdif), bilinear in the four
// cells whose centres surround the gauge when all four are wet, else the
// containing cell's values (the flume port's convention,
// gradient_particle_filter/flume_solver.py:91-99, in two dimensions).  A
// gauge whose cells are not all interior gets NaN.  xy: (x, y) per gauge.
// Fine -> coarse average (update.f90:112-203): one thread per coarse cell of
// the ni x nj block from (ci0, cj0).
__global__ void k_amr_average(double* qc, const double* __restrict__ auxc, int mauxc, int mitc,
                              const double* __restrict__ qf, const double* __restrict__ auxf,
                              int mauxf, int mitf, int ci0, int cj0, int ni, int nj, int fi0,
                              int fj0, int rx, int ry, double dry_tolerance)
{
    const long n = (long)ni * nj;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int a = (int)(t % ni), b = (int)(t / ni);
        amr_average_cell(qc, auxc, mauxc, mitc, qf, auxf, mauxf, mitf, ci0 + a, cj0 + b,
                         fi0 + a * rx, fj0 + b * ry, rx, ry, dry_tolerance);
    }
}

// Coarse -> fine interpolation (filval.f90:190-381): one thread per interior
// cell of the enlarged coarse patch.
__global__ void k_amr_interp(const double* __restrict__ valc, const double* __restrict__ auxc,
                             int mauxc, int mic, int mjc, double* val,
                             const double* __restrict__ auxf, int mauxf, int mitot,
                             const unsigned char* __restrict__ needs_set, int rx, int ry,
                             double dry_tolerance, double sea_level, double rho_f, double grav)
{
    const int ni = mic - 2;
    const long n = (long)ni * (mjc - 2);
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(t % ni), j = 1 + (int)(t / ni);
        amr_interp_cell(valc, auxc, mauxc, mic, val, auxf, mauxf, mitot, needs_set, rx, ry, i, j,
                        dry_tolerance, sea_level, rho_f, grav);
    }
}

// ---- region fill (filpatch.f90's filrecur on device)

__global__ void k_amr_region_init(dclaw_amr_region R, int iregsz, int jregsz)
{
    const int ni = R.ihi - R.ilo + 1;
    const long n = (long)ni * (R.jhi - R.jlo + 1);
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int gi = R.ilo + (int)(t % ni), gj = R.jlo + (int)(t / ni);
        R.flag[t] = amr_in_domain(gi, gj, iregsz, jregsz) ? 0 : 1;
    }
}

// One thread per cell of the intersection ia..ib x ja..jb.
__global__ void k_amr_copy(dclaw_amr_region R, const double* __restrict__ sq,
                           const double* __restrict__ sq_old, double w_old,
                           const double* __restrict__ saux, int smaux, int silo, int sjlo, int smx,
                           int copy_aux, int ia, int ib, int ja, int jb)
{
    const int ni = ib - ia + 1;
    const long n = (long)ni * (jb - ja + 1);
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x)
        amr_copy_cell(R, sq, sq_old, w_old, saux, smaux, silo, sjlo, smx, copy_aux != 0,
                      ia + (int)(t % ni), ja + (int)(t / ni));
}

// trimbd: one workgroup; per-thread extremes of the unset cells, then a tree
// in LDS.
__global__ __launch_bounds__(1024) void k_amr_unset_box(dclaw_amr_region R, int* box)
{
    __shared__ int sh[4][1024];
    const int ni = R.ihi - R.ilo + 1;
    const long n = (long)ni * (R.jhi - R.jlo + 1);
    int lo_i = INT_MAX, hi_i = INT_MIN, lo_j = INT_MAX, hi_j = INT_MIN;
    for (long t = threadIdx.x; t < n; t += 1024) {
        if (R.flag[t] != 0) continue;
        const int gi = R.ilo + (int)(t % ni), gj = R.jlo + (int)(t / ni);
        lo_i = min(lo_i, gi); hi_i = max(hi_i, gi);
        lo_j = min(lo_j, gj); hi_j = max(hi_j, gj);
    }
    const int tid = (int)threadIdx.x;
    sh[0][tid] = lo_i; sh[1][tid] = hi_i; sh[2][tid] = lo_j; sh[3][tid] = hi_j;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (tid < h) {
            sh[0][tid] = min(sh[0][tid], sh[0][tid + h]);
            sh[1][tid] = max(sh[1][tid], sh[1][tid + h]);
            sh[2][tid] = min(sh[2][tid], sh[2][tid + h]);
            sh[3][tid] = max(sh[3][tid], sh[3][tid + h]);
        }
        __syncthreads();
    }
    if (tid < 4) box[tid] = sh[tid][0];
}

// filpatch.f90's interpolation: one thread per interior coarse cell.
__global__ void k_amr_filpatch(dclaw_amr_region R, const double* __restrict__ valc,
                               const double* __restrict__ auxc, AmrFill G)
{
    const int ni = G.mxc - 2;
    const long n = (long)ni * (G.myc - 2);
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x)
        amr_filpatch_cell(valc, auxc, R.q, R.aux, R.flag, G, 2 + (int)(t % ni),
                          2 + (int)(t / ni));
}

__global__ void k_amr_region_bc(dclaw_amr_region R, int iregsz, int jregsz, int b0, int b1,
                                int b2, int b3)
{
    const int bc[4] = {b0, b1, b2, b3};
    const int ni = R.ihi - R.ilo + 1;
    const long n = (long)ni * (R.jhi - R.jlo + 1);
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int gi = R.ilo + (int)(t % ni), gj = R.jlo + (int)(t / ni);
        if (!amr_in_domain(gi, gj, iregsz, jregsz))
            amr_region_bc_cell(R, bc, iregsz, jregsz, gi, gj);
    }
}

// setaux.f90:59-139 (bed of the cells not copied from an old patch), then
// :154-184 (cells outside the domain take the bed of the nearest domain
// cell) in a second launch, then :186-258 (the D-Claw fields).
__global__ void k_amr_setaux_bed(double* aux, int maux, int mx, int my, int ilo, int jlo,
                                 const unsigned char* __restrict__ bed_set,
                                 const double* __restrict__ bed, int iregsz, int jregsz)
{
    const int nx = mx + 4;
    const long n = (long)nx * (my + 4);
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int gi = ilo - 2 + (int)(t % nx), gj = jlo - 2 + (int)(t / nx);
        if (!bed_set[t] && amr_in_domain(gi, gj, iregsz, jregsz)) aux[(size_t)maux * t] = bed[t];
    }
}

__global__ void k_amr_setaux_fields(double* aux, int maux, int mx, int my, int ilo, int jlo,
                                    int iregsz, int jregsz, double phi, double theta)
{
    const int nx = mx + 4;
    const long n = (long)nx * (my + 4);
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int gi = ilo - 2 + (int)(t % nx), gj = jlo - 2 + (int)(t / nx);
        double* a = aux + (size_t)maux * t;
        if (!amr_in_domain(gi, gj, iregsz, jregsz)) {
            const int si = min(max(gi, 0), iregsz - 1), sj = min(max(gj, 0), jregsz - 1);
            a[0] = aux[(size_t)maux * ((size_t)(si - ilo + 2) + (size_t)nx * (size_t)(sj - jlo + 2))];
        }
        for (int k = 1; k < maux; k++) a[k] = 0.0;
        a[1] = phi;
        a[2] = theta;
    }
}

// With dq (the tangent of q) the tangent of each gauge value goes to dout,
// through the same branch and weights as the value (the bed has no tangent).
__global__ void k_gauges(const double* __restrict__ q, const double* __restrict__ aux, int maux,
                         int mx, int my, double xlower, double ylower, double dx, double dy,
                         double drytol, const double* __restrict__ xy, int ng, double* out,
                         const double* __restrict__ dq, double* dout)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const double fx = (xy[2 * g] - xlower) / dx, fy = (xy[2 * g + 1] - ylower) / dy;
    const int ic = (int)floor(fx) + 1, jc = (int)floor(fy) + 1;
    const int i0 = (int)floor(fx + 0.5), j0 = (int)floor(fy + 0.5);
    const double xo = fx + 0.5 - (double)i0, yo = fy + 0.5 - (double)j0;
    double* o = out + 8 * g;
    if (!(fx >= 0.5 && fy >= 0.5 && i0 >= 1 && i0 + 1 <= mx && j0 >= 1 && j0 + 1 <= my)) {
        for (int m = 0; m < 8; m++) o[m] = __longlong_as_double(0x7ff8000000000000LL);
        if (dq)
            for (int m = 0; m < 8; m++) dout[8 * g + m] = __longlong_as_double(0x7ff8000000000000LL);
        return;
    }
    bool wet = true;
    for (int dj = 0; dj < 2; dj++)
        for (int di = 0; di < 2; di++)
            if (!(q[qidx(0, i0 + di, j0 + dj, mx)] > drytol)) wet = false;
    if (wet) {
        const double w00 = (1.0 - xo) * (1.0 - yo), w10 = xo * (1.0 - yo);
        const double w01 = (1.0 - xo) * yo, w11 = xo * yo;
        for (int m = 0; m < 7; m++)
            o[m] = ((w00 * q[qidx(m, i0, j0, mx)] + w10 * q[qidx(m, i0 + 1, j0, mx)]) +
                    w01 * q[qidx(m, i0, j0 + 1, mx)]) +
                   w11 * q[qidx(m, i0 + 1, j0 + 1, mx)];
        double topo = 0.0;
        for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++) {
                const double w = (di ? xo : 1.0 - xo) * (dj ? yo : 1.0 - yo);
                topo = topo + w * (aux[aidx(i0 + di, j0 + dj, mx, maux)] -
                                   q[qidx(6, i0 + di, j0 + dj, mx)]);
            }
        o[7] = o[0] + topo;
        if (dq) {
            double* d = dout + 8 * g;
            for (int m = 0; m < 7; m++)
                d[m] = ((w00 * dq[qidx(m, i0, j0, mx)] + w10 * dq[qidx(m, i0 + 1, j0, mx)]) +
                        w01 * dq[qidx(m, i0, j0 + 1, mx)]) +
                       w11 * dq[qidx(m, i0 + 1, j0 + 1, mx)];
            d[7] = d[0] - d[6];
        }
    } else {
        for (int m = 0; m < 7; m++) o[m] = q[qidx(m, ic, jc, mx)];
        o[7] = o[0] + (aux[aidx(ic, jc, mx, maux)] - q[qidx(6, ic, jc, mx)]);
        if (dq) {
            double* d = dout + 8 * g;
            for (int m = 0; m < 7; m++) d[m] = dq[qidx(m, ic, jc, mx)];
            d[7] = d[0] - d[6];
        }
    }
}


// ---- adjoint step (dclaw_adjoint.h), tile-sparse: the stages run over the
// 16 x 4 cell tiles within ADJ_HALO cells of a wet cell (and the tiles of the
// border band, where the boundary map gathers ghost adjoints), one wavefront
// per tile, tiles taken from a compacted list by a persistent grid.  Outside
// those tiles the step maps a cell to its dry state and every intermediate of
// the stages is 0, so the workspace planes of inactive tiles read as 0 (a
// tile leaving the active set is cleared at the next call), bdif's plane is
// written for every cell, and one dense pass writes the adjoint of the dry
// map (only bdif passes).  The batched form runs the same passes over every
// member's tiles at once (one launch per pass for the whole ensemble).
enum AdjStage {
    ADJ_REC_PREPARE, ADJ_REC_TAUDIR, ADJ_REC_WAVES, ADJ_REC_FLUX, ADJ_REC_SPLIT, ADJ_REC_UPDATE,
    ADJ_REV_SRC2, ADJ_REV_SPLIT, ADJ_REV_FLUX, ADJ_REV_WAVES, ADJ_REV_TAUDIR, ADJ_REV_PREPARE,
    ADJ_REV_BC, ADJ_REV_SRC2_SEEDED, ADJ_UPD_SRC2, ADJ_NSTAGES
};
// Tiles are 16 cells wide and 4 high (64 cells: one wavefront): a tile row
// of a workspace plane is one 128-byte line (the planes' row stride is a
// multiple of 16 doubles), so a tile reads whole lines of its own cells.
constexpr int ADJ_TW = dclaw_adj::TILE_W;
constexpr int ADJ_TH = dclaw_adj::TILE_H;
constexpr int ADJ_HALO = 5;       // cells around a wet cell whose tiles are active
constexpr int ADJ_BLOCK = 256;    // dense passes
constexpr int ADJ_GRID = 8192;    // persistent grid of the tile passes
constexpr int ADJ_GP_PARTS = 256; // partial sums per parameter direction

// The tile passes cut the active list into ADJ_XCDS contiguous chunks and
// give chunk x to the workgroups b with b % ADJ_XCDS == x.  The dispatcher
// deals workgroups to the 8 XCDs round-robin by index, so one XCD works on a
// run of consecutive list entries -- neighbouring tiles, whose halo reads
// share cache lines -- and they meet in that XCD's L2 instead of being
// fetched once per XCD.  Each entry is still taken by exactly one workgroup
// (the grid is a multiple of ADJ_XCDS), so results do not depend on it.
constexpr int ADJ_XCDS = 8;

__device__ __forceinline__ void adj_span(int n, int& k, int& end, int& step)
{
    const int x = (int)(blockIdx.x % ADJ_XCDS);
    step = (int)(gridDim.x / ADJ_XCDS);
    k = (int)((long long)n * x / ADJ_XCDS) + (int)(blockIdx.x / ADJ_XCDS);
    end = (int)((long long)n * (x + 1) / ADJ_XCDS);
}

// Grid of the tile passes: up to ADJ_GRID workgroups, a multiple of ADJ_XCDS.
inline unsigned adj_grid(long long ntiles)
{
    const long long g = std::min<long long>(ntiles, ADJ_GRID);
    return (unsigned)std::max<long long>(ADJ_XCDS, (g + ADJ_XCDS - 1) / ADJ_XCDS * ADJ_XCDS);
}

struct AdjTiles {
    int ntx, nty;
    int* flags;       // ntx * nty: the marks of this call's tiles (zeroed at its start)
    int* list;        // active tiles of the last call, in tile order
    int* count;       // [0]: their number
    int* bcount;      // marked tiles per block of 64 tiles, then their list offsets
};

// A patch's tile bookkeeping at the end of its workspace: after the stage
// planes and the parameter-gradient partial sums come the tile marks, the
// active list and its counter (ints).
__host__ __device__ inline AdjTiles adj_tiles(double* work, int mx, int my, size_t N)
{
    AdjTiles T;
    T.ntx = (mx + 4 + ADJ_TW - 1) / ADJ_TW;
    T.nty = (my + 4 + ADJ_TH - 1) / ADJ_TH;
    const size_t ntiles = (size_t)T.ntx * (size_t)T.nty;
    int* ints = (int*)(work + (size_t)dclaw_adj::W_FIELDS * N +
                       (size_t)dclaw_adj::MAXP * ADJ_GP_PARTS);
    T.flags = ints;
    T.list = ints + ntiles;
    T.count = ints + 2 * ntiles;
    T.bcount = ints + 2 * ntiles + 1;
    return T;
}

// Ordered compaction of marked tiles (the active list comes out in tile
// order, so a run of list entries is a run of neighbouring tiles):
// k_*_count_marks counts the marks of each block of 64 tiles, k_scan_blocks
// turns the counts into list offsets (and writes the total), and the list
// kernel writes each marked tile at its block's offset plus its rank in the
// block.
__global__ __launch_bounds__(1024) void k_scan_blocks(int* a, int n, int* total)
{
    __shared__ int sh[1024];
    const int tid = (int)threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int lo = min(n, tid * per), hi = min(n, lo + per);
    int sum = 0;
    for (int k = lo; k < hi; k++) sum += a[k];
    sh[tid] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = tid >= off ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += v;
        __syncthreads();
    }
    int run = sh[tid] - sum;
    for (int k = lo; k < hi; k++) {
        const int c = a[k];
        a[k] = run;
        run += c;
    }
    if (tid == 1023) *total = sh[1023];
}

template <int STAGE>
__device__ __forceinline__ void adj_stage(const dclaw_adj::AdjArgs& A, const dclaw_adj::TP& Pz,
                                          int i, int j)
{
    if constexpr (STAGE == ADJ_REC_PREPARE) dclaw_adj::rec_prepare(A, Pz, i, j);
    else if constexpr (STAGE == ADJ_REC_TAUDIR) dclaw_adj::rec_taudir(A, Pz, i, j);
    else if constexpr (STAGE == ADJ_REC_WAVES) dclaw_adj::rec_waves(A, Pz, i, j);
    else if constexpr (STAGE == ADJ_REC_FLUX) dclaw_adj::rec_flux(A, Pz, i, j);
    else if constexpr (STAGE == ADJ_REC_SPLIT) dclaw_adj::rec_split(A, Pz, i, j);
    else if constexpr (STAGE == ADJ_REC_UPDATE) dclaw_adj::rec_update(A, Pz, i, j);
    else if constexpr (STAGE == ADJ_REV_SRC2) dclaw_adj::rev_src2_t<true>(A, Pz, i, j);
    else if constexpr (STAGE == ADJ_REV_SRC2_SEEDED) dclaw_adj::rev_src2_t<false>(A, Pz, i, j);
    else if constexpr (STAGE == ADJ_UPD_SRC2) dclaw_adj::upd_src2(A, Pz, i, j);
    else if constexpr (STAGE == ADJ_REV_SPLIT) dclaw_adj::rev_split(A, Pz, i, j);
    else if constexpr (STAGE == ADJ_REV_FLUX) dclaw_adj::rev_flux(A, Pz, i, j);
    else if constexpr (STAGE == ADJ_REV_WAVES) dclaw_adj::rev_waves(A, Pz, i, j);
    else if constexpr (STAGE == ADJ_REV_TAUDIR) dclaw_adj::rev_taudir(A, Pz, i, j);
    else if constexpr (STAGE == ADJ_REV_PREPARE) dclaw_adj::rev_prepare(A, Pz, i, j);
    else dclaw_adj::rev_bc(A, i, j);
}

// Cell (i, j) of tile t for this lane; false outside the grid.
__device__ __forceinline__ bool adj_tile_cell(const dclaw_adj::AdjArgs& A, const AdjTiles& T,
                                              int t, int& i, int& j)
{
    i = -1 + (t % T.ntx) * ADJ_TW + (int)(threadIdx.x % ADJ_TW);
    j = -1 + (t / T.ntx) * ADJ_TH + (int)(threadIdx.x / ADJ_TW);
    return i <= A.mx + 2 && j <= A.my + 2;
}

// Clear every plane of cell n but bdif's (rewritten densely each call).
__device__ __forceinline__ void adj_clear_cell(const dclaw_adj::AdjArgs& A, size_t n)
{
    for (int f = 0; f < dclaw_adj::W_FIELDS; f++)
        if (f != dclaw_adj::W_PC + dclaw_adj::PC_BDIF) A.W[(size_t)f * A.N + n] = 0.0;
}

// Dense work of cell n: bdif's plane (max(bdif, 0) of the loaded state), the
// dry state as q_out (if wanted) and the adjoint of the dry map as lam_in
// (the tile passes overwrite both in active tiles): with the boundary map on
// load a ghost cell gets 0 and an interior cell of the border band also the
// bdif adjoints its dry ghost images send it (the boundary pass's sum, in its
// order); marks the tiles within ADJ_HALO of a wet cell, or every tile when
// dry_tolerance <= 0.
__device__ __forceinline__ void adj_dense_cell(const dclaw_adj::AdjArgs& A, const AdjTiles& T,
                                               size_t n)
{
    const int i = (int)(n % (size_t)A.ld) - 1, j = (int)(n / (size_t)A.ld) - 1;
    if (i > A.mx + 2) return;  // row padding: never read, stays 0
    const double drytol = A.P.dry_tolerance;
    double q7[7], fu, fv;
    dclaw_adj::load_q(A, i, j, q7, fu, fv);
    const double bd = q7[6] > 0.0 ? q7[6] : 0.0;
    A.W[(size_t)(dclaw_adj::W_PC + dclaw_adj::PC_BDIF) * A.N + n] = bd;
    if (A.q_out) {
        double* o = A.q_out + qidx(0, i, j, A.mx);
        for (int m = 0; m < 6; m++) o[m] = 0.0;
        o[6] = bd;
    }
    {
        double l6 = 0.0;
        if (A.bc[0] == 0 || dclaw_adj::interior(A, i, j)) {
            const double* lo = A.lam_out + qidx(0, i, j, A.mx);
            l6 = q7[6] > 0.0 ? lo[6] : 0.0;
            if (dclaw_adj::bc_gathers(A, i, j))
                dclaw_adj::for_ghost_images(A, i, j, [&](int ci, int cj, double, double) {
                    l6 += dclaw_adj::dry_ghost_bdif(A, ci, cj, q7[6]);
                });
        }
        double* li = A.lam_in + qidx(0, i, j, A.mx);
        for (int m = 0; m < 6; m++) li[m] = 0.0;
        li[6] = l6;
    }
    const bool wet = !(drytol > 0.0) || q7[0] > drytol;
    if (wet) {
        const int ia = max(i - ADJ_HALO, -1), ib = min(i + ADJ_HALO, A.mx + 2);
        const int ja = max(j - ADJ_HALO, -1), jb = min(j + ADJ_HALO, A.my + 2);
        for (int ty = (ja + 1) / ADJ_TH; ty <= (jb + 1) / ADJ_TH; ty++)
            for (int tx = (ia + 1) / ADJ_TW; tx <= (ib + 1) / ADJ_TW; tx++)
                T.flags[ty * T.ntx + tx] = 1;
    }
}

// ---- one patch

template <int STAGE>
__global__ __launch_bounds__(64) void k_adjoint(dclaw_adj::AdjArgs A, AdjTiles T)
{
    A.P = *A.Pd;
    A.dt = *A.dt_dev;
    A.tflags = T.flags;
    A.tntx = T.ntx;
    const dclaw_adj::TP Pz = dclaw_adj::zero_params(A.P);
    int k, end, step;
    adj_span(*T.count, k, end, step);
    for (; k < end; k += step) {
        int i, j;
        if (adj_tile_cell(A, T, T.list[k], i, j)) adj_stage<STAGE>(A, Pz, i, j);
    }
}

// The same, held to WAVES waves per SIMD (the stages the batched step holds
// there too).
template <int STAGE, int WAVES>
__global__ __launch_bounds__(64, WAVES) void k_adjointw(dclaw_adj::AdjArgs A, AdjTiles T)
{
    A.P = *A.Pd;
    A.dt = *A.dt_dev;
    A.tflags = T.flags;
    A.tntx = T.ntx;
    const dclaw_adj::TP Pz = dclaw_adj::zero_params(A.P);
    int k, end, step;
    adj_span(*T.count, k, end, step);
    for (; k < end; k += step) {
        int i, j;
        if (adj_tile_cell(A, T, T.list[k], i, j)) adj_stage<STAGE>(A, Pz, i, j);
    }
}

// Clear the tiles of the previous call's list that are not marked for this
// call (every tile pass rewrites all planes of its own cells, so the tiles
// that stay active need no clearing).
__global__ __launch_bounds__(64) void k_adjoint_clear_stale(dclaw_adj::AdjArgs A, AdjTiles T)
{
    const int nold = *T.count;
    for (int k = blockIdx.x; k < nold; k += gridDim.x) {
        const int t = T.list[k];
        int i, j;
        if (T.flags[t] == 0 && adj_tile_cell(A, T, t, i, j))
            adj_clear_cell(A, dclaw_adj::cix(A, i, j));
    }
}

__global__ __launch_bounds__(ADJ_BLOCK) void k_adjoint_dense(dclaw_adj::AdjArgs A, AdjTiles T)
{
    const size_t n = (size_t)blockIdx.x * ADJ_BLOCK + threadIdx.x;
    if (n >= A.N) return;
    A.P = *A.Pd;
    adj_dense_cell(A, T, n);
}

// With the boundary map on load, a tile of the border band (holding ghost
// cells or interior cells within 2 of the edge) next to a tile marked for a
// wet cell is marked too (2): the boundary pass of a cell gathers its ghost
// images' adjoints, which lie in its own tile or a neighbour, and a tile the
// passes skip has them from the dense pass, which knows only the dry ones.
// A tile marked 2 is beyond ADJ_HALO of every wet cell, so its ghosts' values
// are the dry ones too.
__device__ __forceinline__ void adj_mark_border(const AdjTiles& T, int mx, int my, int t)
{
    const int tx = t % T.ntx, ty = t / T.ntx;
    const bool bx = ADJ_TW * tx <= 3 || ADJ_TW * tx + ADJ_TW - 1 >= mx;   // i + 1 <= 3, >= mx
    const bool by = ADJ_TH * ty <= 3 || ADJ_TH * ty + ADJ_TH - 1 >= my;
    if (!(bx || by) || T.flags[t] != 0) return;
    bool near = false;
    for (int y = max(0, ty - 1); y <= min(T.nty - 1, ty + 1); y++)
        for (int x = max(0, tx - 1); x <= min(T.ntx - 1, tx + 1); x++)
            near = near || T.flags[y * T.ntx + x] == 1;
    if (near) T.flags[t] = 2;
}

__global__ __launch_bounds__(256) void k_adjoint_border(AdjTiles T, int mx, int my)
{
    const int t = (int)(blockIdx.x * 256 + threadIdx.x);
    if (t < T.ntx * T.nty) adj_mark_border(T, mx, my, t);
}

// Marked tiles per block of 64 (pass 1 of the ordered compaction).
__global__ __launch_bounds__(64) void k_adjoint_count_marks(AdjTiles T, int ntiles)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    const unsigned long long b = __ballot(t < ntiles && T.flags[t] != 0);
    if (threadIdx.x == 0) T.bcount[blockIdx.x] = (int)__popcll(b);
}

// Write the marked tiles into the list in tile order (T.bcount holds each
// block's offset); the marks stay for the boundary pass.  Every tile pass writes only its own
// cells, so the order only affects cache locality.
__global__ __launch_bounds__(64) void k_adjoint_list(AdjTiles T, int ntiles)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    const bool f = t < ntiles && T.flags[t] != 0;
    const unsigned long long b = __ballot(f);
    if (f) {
        const unsigned long long below = b & ((1ull << threadIdx.x) - 1ull);
        T.list[T.bcount[blockIdx.x] + (int)__popcll(below)] = t;
    }
}

// Sum of the per-cell parameter gradients, in a fixed order: block (b, p)
// sums chunk b of direction p's plane (strided partial sums, then a tree in
// LDS) into part[p][b]; adj_gp_final adds the ADJ_GP_PARTS partials of each
// direction the same way.
__device__ __forceinline__ void adj_gp_part(const double* __restrict__ W, size_t N, int p, int b,
                                            double* __restrict__ part)
{
    __shared__ double sh[256];
    const double* g = W + (size_t)(dclaw_adj::W_GP + p) * N;
    const size_t chunk = (N + ADJ_GP_PARTS - 1) / ADJ_GP_PARTS;
    const size_t lo = (size_t)b * chunk;
    const size_t hi = lo + chunk < N ? lo + chunk : N;
    // four independent sums per thread (four loads in flight), combined in a
    // fixed order
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    size_t n = lo + threadIdx.x;
    for (; n + 768 < hi; n += 1024) {
        s0 += g[n];
        s1 += g[n + 256];
        s2 += g[n + 512];
        s3 += g[n + 768];
    }
    for (; n < hi; n += 256) s0 += g[n];
    sh[threadIdx.x] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[p * ADJ_GP_PARTS + b] = sh[0];
}

__device__ __forceinline__ void adj_gp_final(const double* __restrict__ part, int p,
                                             double* __restrict__ out)
{
    __shared__ double sh[ADJ_GP_PARTS];
    sh[threadIdx.x] = part[p * ADJ_GP_PARTS + threadIdx.x];
    __syncthreads();
    for (int h = ADJ_GP_PARTS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[p] = sh[0];
}

__global__ __launch_bounds__(256) void k_adjoint_gp(const double* __restrict__ W, size_t N,
                                                    double* __restrict__ part)
{
    adj_gp_part(W, N, blockIdx.y, blockIdx.x, part);
}

__global__ __launch_bounds__(ADJ_GP_PARTS) void k_adjoint_gp_final(const double* __restrict__ part,
                                                                   double* __restrict__ out)
{
    adj_gp_final(part, blockIdx.x, out);
}

// ---- many patches (one launch per pass for the whole ensemble)

struct AdjBatch {
    const dclaw_adj_patch* pt;
    int npatch;
    int bc[4];
    const dclaw_params* params;    // per member
    const dclaw_params* dparams;   // np per member
    int np;
    const double* dt;              // per member
    long long* cell_prefix;        // npatch + 1
    long long* tile_prefix;        // npatch + 1
    int* glist;                    // active tiles in global tile order: the tile within its patch
    int* gmem;                     // and its patch (parallel to glist)
    int* gcount;
    int* bcount;                   // marks per block of 64 global tiles, then list offsets
};

// Largest m with prefix[m] <= g.
__device__ __forceinline__ int adj_find(const long long* prefix, int npatch, long long g)
{
    int lo = 0, hi = npatch - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (prefix[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ dclaw_adj::AdjArgs adj_member(const AdjBatch& B, int m, AdjTiles& T)
{
    const dclaw_adj_patch& p = B.pt[m];
    dclaw_adj::AdjArgs A;
    A.q_in = p.q_in;
    A.aux = p.aux;
    A.maux = p.maux;
    A.mx = p.mx;
    A.my = p.my;
    A.dx = p.dx;
    A.dy = p.dy;
    for (int k = 0; k < 4; k++) A.bc[k] = B.bc[k];
    A.Pd = B.params + p.member;
    A.dt_dev = B.dt + p.member;
    A.P = *A.Pd;
    A.dt = *A.dt_dev;
    A.dP = B.dparams ? B.dparams + (size_t)p.member * (size_t)B.np : nullptr;
    A.np = B.np;
    A.lam_out = p.lam_out;
    A.lam_in = p.lam_in;
    A.q_out = p.q_out;
    A.W = p.work;
    A.N = dclaw_adj::plane_size(p.mx, p.my);
    A.ld = dclaw_adj::plane_ld(p.mx);
    T = adj_tiles(p.work, p.mx, p.my, A.N);
    A.tflags = T.flags;
    A.tntx = T.ntx;
    return A;
}

// Cell and tile prefixes of the members: one block, 1024 members at a time,
// an inclusive scan in LDS plus the running carry (integer sums: the result
// does not depend on the order).
__global__ __launch_bounds__(1024) void k_adjb_prefix(AdjBatch B)
{
    __shared__ long long sc[1024], st[1024];
    const int tid = (int)threadIdx.x;
    long long carry_c = 0, carry_t = 0;
    for (int base = 0; base < B.npatch; base += 1024) {
        const int m = base + tid;
        long long c = 0, t = 0;
        if (m < B.npatch) {
            const int mx = B.pt[m].mx, my = B.pt[m].my;
            c = (long long)dclaw_adj::plane_size(mx, my);
            t = (long long)((mx + 4 + ADJ_TW - 1) / ADJ_TW) * ((my + 4 + ADJ_TH - 1) / ADJ_TH);
        }
        sc[tid] = c;
        st[tid] = t;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const long long vc = tid >= off ? sc[tid - off] : 0;
            const long long vt = tid >= off ? st[tid - off] : 0;
            __syncthreads();
            sc[tid] += vc;
            st[tid] += vt;
            __syncthreads();
        }
        if (m < B.npatch) {
            B.cell_prefix[m] = carry_c + sc[tid] - c;
            B.tile_prefix[m] = carry_t + st[tid] - t;
        }
        carry_c += sc[1023];
        carry_t += st[1023];
        __syncthreads();
    }
    if (tid == 0) {
        B.cell_prefix[B.npatch] = carry_c;
        B.tile_prefix[B.npatch] = carry_t;
    }
    // gcount[1]: whether some member's src2 reverse takes the seeded products
    __shared__ int seeded;
    if (tid == 0) seeded = 0;
    __syncthreads();
    for (int m = tid; m < B.npatch; m += 1024)
        if (!dclaw_adj::src2_vjp_covers(B.params[B.pt[m].member])) seeded = 1;
    __syncthreads();
    if (tid == 0) B.gcount[1] = seeded;
}

__global__ __launch_bounds__(ADJ_BLOCK) void k_adjb_dense(AdjBatch B, long long ncells)
{
    const long long g = (long long)blockIdx.x * ADJ_BLOCK + threadIdx.x;
    if (g >= ncells) return;
    const int m = adj_find(B.cell_prefix, B.npatch, g);
    AdjTiles T;
    const dclaw_adj::AdjArgs A = adj_member(B, m, T);
    adj_dense_cell(A, T, (size_t)(g - B.cell_prefix[m]));
}

__global__ __launch_bounds__(64) void k_adjb_clear_stale(AdjBatch B)
{
    AdjTiles T;
    const dclaw_adj::AdjArgs A = adj_member(B, blockIdx.y, T);
    const int nold = *T.count;
    for (int k = blockIdx.x; k < nold; k += gridDim.x) {
        const int t = T.list[k];
        int i, j;
        if (T.flags[t] == 0 && adj_tile_cell(A, T, t, i, j))
            adj_clear_cell(A, dclaw_adj::cix(A, i, j));
    }
}

// Zero member blockIdx.y's tile marks (start of a call).
__global__ __launch_bounds__(256) void k_adjb_zero_flags(AdjBatch B)
{
    const dclaw_adj_patch& p = B.pt[blockIdx.y];
    const AdjTiles T = adj_tiles(p.work, p.mx, p.my, dclaw_adj::plane_size(p.mx, p.my));
    for (int t = (int)(blockIdx.x * 256 + threadIdx.x); t < T.ntx * T.nty;
         t += (int)(gridDim.x * 256))
        T.flags[t] = 0;
}

__global__ __launch_bounds__(256) void k_adjb_border(AdjBatch B)
{
    const dclaw_adj_patch& p = B.pt[blockIdx.y];
    const AdjTiles T = adj_tiles(p.work, p.mx, p.my, dclaw_adj::plane_size(p.mx, p.my));
    for (int t = (int)(blockIdx.x * 256 + threadIdx.x); t < T.ntx * T.nty;
         t += (int)(gridDim.x * 256))
        adj_mark_border(T, p.mx, p.my, t);
}

__global__ __launch_bounds__(256) void k_adjb_count_zero(AdjBatch B)
{
    const int m = blockIdx.x * 256 + threadIdx.x;
    if (m == 0) *B.gcount = 0;
    if (m >= B.npatch) return;
    const dclaw_adj_patch& p = B.pt[m];
    adj_tiles(p.work, p.mx, p.my, dclaw_adj::plane_size(p.mx, p.my)).count[0] = 0;
}

// Whether global tile g (< ntiles) is marked, and its patch's bookkeeping.
__device__ __forceinline__ bool adjb_marked(const AdjBatch& B, long long g, AdjTiles& T, int& t)
{
    const int m = adj_find(B.tile_prefix, B.npatch, g);
    const dclaw_adj_patch& p = B.pt[m];
    T = adj_tiles(p.work, p.mx, p.my, dclaw_adj::plane_size(p.mx, p.my));
    t = (int)(g - B.tile_prefix[m]);
    return T.flags[t] != 0;
}

// Marked tiles per block of 64 global tiles (pass 1 of the ordered
// compaction).
__global__ __launch_bounds__(64) void k_adjb_count_marks(AdjBatch B, long long ntiles)
{
    const long long g = (long long)blockIdx.x * 64 + threadIdx.x;
    AdjTiles T;
    int t = 0;
    const bool f = g < ntiles && adjb_marked(B, g, T, t);
    const unsigned long long b = __ballot(f);
    if (threadIdx.x == 0) B.bcount[blockIdx.x] = (int)__popcll(b);
}

// Compaction over every patch's marks: each marked tile goes to its patch's
// own list (for the next call's clearing; any order) and to the global list,
// in global tile order (B.bcount holds each block's offset).  The lanes whose
// tiles belong to the wave's first marked patch (all of them but at a
// patch boundary) take their slots of that patch's list with one atomic.
__global__ __launch_bounds__(64) void k_adjb_list(AdjBatch B, long long ntiles)
{
    const long long g = (long long)blockIdx.x * 64 + threadIdx.x;
    AdjTiles T;
    int t = 0, m = -1;
    bool f = false;
    if (g < ntiles) {
        m = adj_find(B.tile_prefix, B.npatch, g);
        const dclaw_adj_patch& p = B.pt[m];
        T = adj_tiles(p.work, p.mx, p.my, dclaw_adj::plane_size(p.mx, p.my));
        t = (int)(g - B.tile_prefix[m]);
        f = T.flags[t] != 0;
    }
    const unsigned long long b = __ballot(f);
    if (b == 0ull) return;
    const int lead = __ffsll((long long)b) - 1;
    const int m0 = __shfl(m, lead);
    const bool agg = f && m == m0;
    const unsigned long long ba = __ballot(agg);
    const unsigned long long below = (1ull << threadIdx.x) - 1ull;
    int base = 0;
    if ((int)threadIdx.x == lead) base = atomicAdd(T.count, (int)__popcll(ba));
    base = __shfl(base, lead);
    if (!f) return;
    if (agg) T.list[base + (int)__popcll(ba & below)] = t;
    else T.list[atomicAdd(T.count, 1)] = t;
    const int slot = B.bcount[blockIdx.x] + (int)__popcll(b & below);
    B.glist[slot] = t;
    B.gmem[slot] = m;
}

// The tile loop of a batched stage pass.  Each list entry carries its patch,
// and the next entry is requested before the current tile's work, so a tile
// starts without a chain of dependent lookups.
template <int STAGE>
__device__ __forceinline__ void adjb_tiles(const AdjBatch& B)
{
    int k, end, step;
    adj_span(*B.gcount, k, end, step);
    if (k >= end) return;
    int t = B.glist[k], m = B.gmem[k];
    for (; k < end; k += step) {
        const int kn = k + step;
        int tn = 0, mn = 0;
        if (kn < end) {
            tn = B.glist[kn];
            mn = B.gmem[kn];
        }
        AdjTiles T;
        const dclaw_adj::AdjArgs A = adj_member(B, m, T);
        const dclaw_adj::TP Pz = dclaw_adj::zero_params(A.P);
        int i, j;
        if (adj_tile_cell(A, T, t, i, j)) adj_stage<STAGE>(A, Pz, i, j);
        t = tn;
        m = mn;
    }
}

template <int STAGE>
__global__ __launch_bounds__(64) void k_adjb(AdjBatch B)
{
    if constexpr (STAGE == ADJ_REV_SRC2_SEEDED)
        if (B.gcount[1] == 0) return;
    adjb_tiles<STAGE>(B);
}

// The same, held to WAVES waves per SIMD (for the stages whose batched form
// needs just over 1/WAVES of the register file).
template <int STAGE, int WAVES>
__global__ __launch_bounds__(64, WAVES) void k_adjbw(AdjBatch B)
{
    adjb_tiles<STAGE>(B);
}

__global__ __launch_bounds__(256) void k_adjb_gp(AdjBatch B)
{
    const dclaw_adj_patch& p = B.pt[blockIdx.z];
    const size_t N = dclaw_adj::plane_size(p.mx, p.my);
    adj_gp_part(p.work, N, blockIdx.y, blockIdx.x, p.work + (size_t)dclaw_adj::W_FIELDS * N);
}

__global__ __launch_bounds__(ADJ_GP_PARTS) void k_adjb_gp_final(AdjBatch B)
{
    const dclaw_adj_patch& p = B.pt[blockIdx.y];
    const size_t N = dclaw_adj::plane_size(p.mx, p.my);
    adj_gp_final(p.work + (size_t)dclaw_adj::W_FIELDS * N, blockIdx.x, p.grad);
}

// Adjoint of the gauge samples (k_gauges): lam_q += d(out)/dq^T lam_out for
// every gauge, gauges in order (one thread, so gauges sharing a cell add in a
// fixed order).  Gauges outside the patch have no derivative.
__global__ void k_gauges_adj(const double* __restrict__ q, const double* __restrict__ aux,
                             int maux, int mx, int my, double xlower, double ylower, double dx,
                             double dy, double drytol, const double* __restrict__ xy, int ng,
                             const double* __restrict__ lam_g, double* lam_q)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int g = 0; g < ng; g++) {
        const double fx = (xy[2 * g] - xlower) / dx, fy = (xy[2 * g + 1] - ylower) / dy;
        const int ic = (int)floor(fx) + 1, jc = (int)floor(fy) + 1;
        const int i0 = (int)floor(fx + 0.5), j0 = (int)floor(fy + 0.5);
        const double xo = fx + 0.5 - (double)i0, yo = fy + 0.5 - (double)j0;
        if (!(fx >= 0.5 && fy >= 0.5 && i0 >= 1 && i0 + 1 <= mx && j0 >= 1 && j0 + 1 <= my))
            continue;
        const double* l = lam_g + 8 * g;
        bool wet = true;
        for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++)
                if (!(q[qidx(0, i0 + di, j0 + dj, mx)] > drytol)) wet = false;
        if (wet) {
            // out[m] = sum_c w_c q(m, c) (m < 7); out[7] = out[0] + sum_c w_c (b_c - q(6, c))
            for (int dj = 0; dj < 2; dj++)
                for (int di = 0; di < 2; di++) {
                    const double wgt = (di ? xo : 1.0 - xo) * (dj ? yo : 1.0 - yo);
                    double* o = lam_q + qidx(0, i0 + di, j0 + dj, mx);
                    for (int m = 0; m < 7; m++) o[m] += wgt * l[m];
                    o[0] += wgt * l[7];
                    o[6] -= wgt * l[7];
                }
        } else {
            double* o = lam_q + qidx(0, ic, jc, mx);
            for (int m = 0; m < 7; m++) o[m] += l[m];
            o[0] += l[7];
            o[6] -= l[7];
        }
    }
    (void)aux;
    (void)maux;
}

int check_patch(int mx, int my, int maux)
{
    if (mx < 1 || my < 1) return fail(DCLAW_EINVAL, "mx and my must be >= 1");
    if (maux < 1) return fail(DCLAW_EINVAL, "maux must be >= 1");
    if ((double)(mx + 4) * (double)(my + 4) * 7.0 > 9.0e18)
        return fail(DCLAW_EINVAL, "patch too large");
    return DCLAW_OK;
}

// Host drop-in state.  The parameters are process-wide configuration (the
// reference's module globals: dclaw_set_params after set_dig); each calling
// thread has its own device buffers and HIP stream (HostCtx), so threads
// that step different patches at once (AMRClaw advanc's OpenMP loop over
// patches) run concurrently.  dclaw_set_params waits for calls in progress.
struct HostCfg {
    std::shared_mutex mu;
    dclaw_params params;
    bool have_params = false;
    dclaw_params* d_params = nullptr;
};
HostCfg g_cfg;

struct HostCtx {
    hipStream_t st = nullptr;
    double* d_scalars = nullptr;   // [0] dt, [1] cfl
    double* d_qa = nullptr;
    double* d_qb = nullptr;
    double* d_aux = nullptr;
    size_t cap_qa = 0, cap_qb = 0, cap_aux = 0;
    double* d_flux[4] = {nullptr, nullptr, nullptr, nullptr};   // step2_ fm, fp, gm, gp
    size_t cap_flux[4] = {0, 0, 0, 0};
    double* d_qc = nullptr;                                      // b4step2 copy for the fluxes
    size_t cap_qc = 0;
    double* d_curv = nullptr;                                    // curvature workspace
    size_t cap_curv = 0;
    int* d_nan = nullptr;                                        // check4nans result
};

// The calling thread's context, created on its first call and kept until the
// process ends (a thread pool keeps its threads; the registry holds them).
int host_ctx(HostCtx** out)
{
    static std::mutex reg_mu;
    static std::vector<HostCtx*> reg;
    thread_local HostCtx* mine = nullptr;
    if (!mine) {
        HostCtx* c = new HostCtx;
        HIP_TRY(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
        HIP_TRY(hipMalloc(&c->d_scalars, 2 * sizeof(double)));
        HIP_TRY(hipMalloc(&c->d_nan, sizeof(int)));
        std::lock_guard<std::mutex> lk(reg_mu);
        reg.push_back(c);
        mine = c;
    }
    *out = mine;
    return DCLAW_OK;
}

int grow(double** p, size_t* cap, size_t need)
{
    if (*cap >= need) return DCLAW_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    HIP_TRY(hipMalloc(p, need * sizeof(double)));
    *cap = need;
    return DCLAW_OK;
}

int validate_params(const dclaw_params& P)
{
    if (P.mwaves != 5) return fail(DCLAW_EINVAL, "mwaves must be 5 (rpn2_dclaw.f:276-315)");
    if (P.order != 1 && P.order != 2) return fail(DCLAW_EINVAL, "order must be 1 or 2");
    if (P.transverse < 0 || P.transverse > 2) return fail(DCLAW_EINVAL, "transverse must be 0..2");
    if (P.src2method < -1 || P.src2method > 2)
        return fail(DCLAW_EINVAL, "src2method must be -1, 0, 1 or 2");
    if (P.src2method == 2 && P.alphamethod != 1)
        return fail(DCLAW_EINVAL, "src2method 2 requires alphamethod 1 (data.py:362-365)");
    if (P.curvature < 0 || P.curvature > 1) return fail(DCLAW_EINVAL, "curvature must be 0 or 1");
    if (P.curvature == 1 && P.entrainment == 1)
        return fail(DCLAW_EINVAL, "curvature with entrainment is not supported (src2 would read "
                                  "beds it eroded earlier in the same sweep)");
    if (P.entrainment < 0 || P.entrainment > 1 || P.entrainment_method < 0 ||
        P.entrainment_method > 1)
        return fail(DCLAW_EINVAL, "entrainment and entrainment_method must be 0 or 1");
    if (!(P.dry_tolerance >= 0.0) || !(P.grav > 0.0))
        return fail(DCLAW_EINVAL, "grav must be > 0 and dry_tolerance >= 0");
    return DCLAW_OK;
}

}  // namespace

extern "C" {

const char* dclaw_last_error(void) { return g_err.c_str(); }

int dclaw_version(void) { return 1; }

int dclaw_strip_width(void) { return STRIP_W; }

int dclaw_set_seg_rows(int seg_rows)
{
    if (seg_rows < 1 || seg_rows > 65536) return fail(DCLAW_EINVAL, "seg_rows must be in [1, 65536]");
    return g_seg_rows.exchange(seg_rows);
}

int dclaw_bc_fill(double* q, int mx, int my, const int* bc, void* stream)
{
    int rc = check_patch(mx, my, 1);
    if (rc) return rc;
    if (!q || !bc) return fail(DCLAW_EINVAL, "null pointer");
    for (int k = 0; k < 4; k++)
        if (bc[k] != DCLAW_BC_EXTRAP && bc[k] != DCLAW_BC_WALL)
            return fail(DCLAW_EINVAL, "bc codes must be 1 (extrap) or 3 (wall)");
    const long n = (long)(mx + 4) * (my + 4);
    const int blocks = (int)std::min<long>((n + 255) / 256, 65535);
    hipLaunchKernelGGL(k_bc_fill, dim3(blocks), dim3(256), 0, (hipStream_t)stream, q, mx, my,
                       bc[0], bc[1], bc[2], bc[3]);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

static int check_bc(const int* bc)
{
    if (!bc) return DCLAW_OK;
    for (int k = 0; k < 4; k++)
        if (bc[k] != DCLAW_BC_EXTRAP && bc[k] != DCLAW_BC_WALL)
            return fail(DCLAW_EINVAL, "bc codes must be 1 (extrap) or 3 (wall)");
    return DCLAW_OK;
}

static int launch_stepgrid(const Tuning& tn, const double* q_in, double* q_out,
                           const double* aux, double* aux_out, int maux, int mx, int my,
                           double dx, double dy, const double* dt_dev,
                           const dclaw_params* params_dev, double* cfl_dev, const int* bc,
                           hipStream_t stream, int* tile_flags = nullptr, int n_flags = 0)
{
    if (int rcb = check_bc(bc)) return rcb;
    int rc = check_patch(mx, my, maux);
    if (rc) return rc;
    if (!q_in || !q_out || !aux || !dt_dev || !params_dev || !cfl_dev)
        return fail(DCLAW_EINVAL, "null pointer");
    if (q_in == q_out) return fail(DCLAW_EINVAL, "q_out must not alias q_in");
    if (!(dx > 0.0) || !(dy > 0.0)) return fail(DCLAW_EINVAL, "dx and dy must be > 0");
    const int nstrips = (mx + STRIP_W - 1) / STRIP_W;
    const int seg_rows = tn.seg_rows;
    const int nseg = (my + seg_rows - 1) / seg_rows;
    if ((long)nstrips * nseg > 0x7fffffffL) return fail(DCLAW_EINVAL, "patch too large for one launch");
    const int b0 = bc ? bc[0] : 0, b1 = bc ? bc[1] : 0, b2 = bc ? bc[2] : 0, b3 = bc ? bc[3] : 0;
    if (!tile_flags) {
        hipLaunchKernelGGL(k_stepgrid<MODE_STEP>, dim3(nstrips * nseg), dim3(64), 0, stream,
                           q_in, q_out, aux, aux_out, maux, mx, my, dx, dy, dt_dev, params_dev,
                           cfl_dev, seg_rows, nstrips, nseg, b0, b1, b2, b3, nullptr, nullptr,
                           nullptr);
        HIP_TRY(hipGetLastError());
        return DCLAW_OK;
    }
    const int ntiles = nstrips * nseg;
    if (n_flags < ntiles) return fail(DCLAW_EINVAL, "tile_flags holds fewer than dclaw_tile_count() ints");
    if (aux_out) return fail(DCLAW_EINVAL, "the split step does not write aux");
    const bool resident = (long)n_flags >= split_resident_ints(ntiles);
    if (resident && !(tn.split_mode == 1 && b0 && b1 && b2 && b3)) {
        // another mode (or stored ghosts) writes q_out without the dry-tile
        // bookkeeping: the next resident call starts over
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, stream, tile_flags + 3 * ntiles + 4,
                           SKIP_META);
    }
    if ((long)n_flags >= 3L * ntiles + 4 && tn.split_mode == 3) {
        // one strip launch: the previous step's wet segments singly and first,
        // the other segments in tall dry groups
        int* wet_list = tile_flags + ntiles;
        int* dry_list = tile_flags + 2 * ntiles;
        int* counters = tile_flags + 3 * ntiles;
        const int ngroups = nstrips * ((nseg + HYB_GROUP - 1) / HYB_GROUP);
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, stream, counters, 2);
        hipLaunchKernelGGL(k_order_hybrid, dim3((ngroups + 255) / 256), dim3(256), 0, stream,
                           tile_flags, nstrips, nseg, wet_list, dry_list, counters);
        hipLaunchKernelGGL(k_stepgrid_hybrid, dim3(ntiles), dim3(64), 0, stream, q_in, q_out, aux,
                           maux, mx, my, dx, dy, dt_dev, params_dev, cfl_dev, seg_rows, nstrips,
                           nseg, b0, b1, b2, b3, tile_flags, wet_list, dry_list, counters);
        HIP_TRY(hipGetLastError());
        const long ntask = (long)ntiles * ((STRIP_W * seg_rows + 63) / 64);
        hipLaunchKernelGGL(k_src2_tiles, dim3((int)std::min<long>(ntask, 16384)), dim3(64), 0,
                           stream, q_out, aux, maux, mx, my, seg_rows, nstrips, ntiles,
                           tile_flags, dt_dev, params_dev);
        HIP_TRY(hipGetLastError());
        return DCLAW_OK;
    }
    if ((long)n_flags >= 2L * ntiles + 4 && tn.split_mode == 1) {
        // one strip launch with the previous step's wet tiles first
        int* order = tile_flags + ntiles;
        int* counters = tile_flags + 2 * ntiles + 2;
        const int* dry_age = nullptr;
        if (resident && b0 && b1 && b2 && b3) {
            // resident workspace, boundary conditions on load: tiles that
            // stayed dry are not read or written
            int* meta = tile_flags + 3 * ntiles + 4;
            int* age = meta + SKIP_META;
            hipLaunchKernelGGL(k_skip_meta, dim3(1), dim3(64), 0, stream, meta, q_in, q_out,
                               seg_rows, nstrips, nseg);
            hipLaunchKernelGGL(k_dry_age, dim3((ntiles + 255) / 256), dim3(256), 0, stream,
                               tile_flags, age, meta, nstrips, nseg,
                               (STRIP_HALO + seg_rows - 1) / seg_rows);
            dry_age = age;
        }
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, stream, counters, 2);
        hipLaunchKernelGGL(k_order_tiles, dim3((ntiles + 255) / 256), dim3(256), 0, stream,
                           tile_flags, ntiles, order, counters);
        hipLaunchKernelGGL(k_stepgrid<MODE_STEP_NOSRC2>, dim3(nstrips * nseg), dim3(64), 0,
                           stream, q_in, q_out, aux, nullptr, maux, mx, my, dx, dy, dt_dev,
                           params_dev, cfl_dev, seg_rows, nstrips, nseg, b0, b1, b2, b3,
                           tile_flags, order, dry_age);
        HIP_TRY(hipGetLastError());
        const long ntask = (long)ntiles * ((STRIP_W * seg_rows + 63) / 64);
        hipLaunchKernelGGL(k_src2_tiles, dim3((int)std::min<long>(ntask, 16384)), dim3(64), 0,
                           stream, q_out, aux, maux, mx, my, seg_rows, nstrips, ntiles,
                           tile_flags, dt_dev, params_dev);
        HIP_TRY(hipGetLastError());
        return DCLAW_OK;
    }
    if ((long)n_flags >= 2L * ntiles + 2 && tn.split_mode == 2) {
        // dry pass, strip body over the wet list, src2 over the wet list
        int* wet_list = tile_flags + ntiles;
        int* counters = tile_flags + 2 * ntiles;
        const int group = std::max(1, std::min(32, tn.dry_rows / seg_rows));
        const int ngroups = (nseg + group - 1) / group;
        const int ncu = cu_count();
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, stream, counters, 2);
        hipLaunchKernelGGL(k_classify, dim3(nstrips * ngroups), dim3(64), 0, stream, q_in, q_out,
                           mx, my, params_dev, seg_rows, group, nseg, nstrips, b0, b1, b2, b3,
                           tile_flags, wet_list, counters);
        hipLaunchKernelGGL(k_stepgrid_list<MODE_STEP_NOSRC2>, dim3(std::min(ntiles, 4 * ncu)),
                           dim3(64), 0, stream, q_in, q_out, aux, maux, mx, my, dx, dy, dt_dev,
                           params_dev, cfl_dev, seg_rows, nstrips, b0, b1, b2, b3, tile_flags,
                           wet_list, counters);
        const long nt = (long)ntiles * ((STRIP_W * seg_rows + 63) / 64);
        hipLaunchKernelGGL(k_src2_list, dim3((int)std::min<long>(nt, 16L * ncu)), dim3(64), 0,
                           stream, q_out, aux, maux, mx, my, seg_rows, nstrips, tile_flags,
                           wet_list, counters, dt_dev, params_dev);
        HIP_TRY(hipGetLastError());
        return DCLAW_OK;
    }
    hipLaunchKernelGGL(k_stepgrid<MODE_STEP_NOSRC2>, dim3(nstrips * nseg), dim3(64), 0, stream,
                       q_in, q_out, aux, nullptr, maux, mx, my, dx, dy, dt_dev, params_dev,
                       cfl_dev, seg_rows, nstrips, nseg, b0, b1, b2, b3, tile_flags, nullptr,
                       nullptr);
    HIP_TRY(hipGetLastError());
    const long ntask = (long)ntiles * ((STRIP_W * seg_rows + 63) / 64);
    hipLaunchKernelGGL(k_src2_tiles, dim3((int)std::min<long>(ntask, 16384)), dim3(64), 0, stream,
                       q_out, aux, maux, mx, my, seg_rows, nstrips, ntiles, tile_flags, dt_dev,
                       params_dev);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_stepgrid_dev(const double* q_in, double* q_out, const double* aux, int maux, int mx,
                       int my, double dx, double dy, const double* dt_dev,
                       const dclaw_params* params_dev, double* cfl_dev, void* stream)
{
    return launch_stepgrid(global_tuning(), q_in, q_out, aux, nullptr, maux, mx, my, dx, dy,
                           dt_dev, params_dev, cfl_dev, nullptr, (hipStream_t)stream);
}

int dclaw_stepgrid_bc_dev(const double* q_in, double* q_out, const double* aux, int maux,
                          int mx, int my, double dx, double dy, const int* bc,
                          const double* dt_dev, const dclaw_params* params_dev,
                          double* cfl_dev, void* stream)
{
    if (!bc) return fail(DCLAW_EINVAL, "null bc");
    return launch_stepgrid(global_tuning(), q_in, q_out, aux, nullptr, maux, mx, my, dx, dy,
                           dt_dev, params_dev, cfl_dev, bc, (hipStream_t)stream);
}

int dclaw_stepgrid_tl_dev(const double* q_in, const double* dq_in, double* q_out, double* dq_out,
                          const double* aux, int maux, int mx, int my, double dx, double dy,
                          const int* bc, const double* dt_dev, const dclaw_params* params_dev,
                          const dclaw_params* dparams_dev, double* cfl_dev, void* stream)
{
    if (int rcb = check_bc(bc)) return rcb;
    if (int rc = check_patch(mx, my, maux)) return rc;
    if (!q_in || !dq_in || !q_out || !dq_out || !aux || !dt_dev || !params_dev || !dparams_dev ||
        !cfl_dev)
        return fail(DCLAW_EINVAL, "null pointer");
    if (q_in == q_out || dq_in == dq_out || q_out == dq_out || q_in == dq_out || dq_in == q_out)
        return fail(DCLAW_EINVAL, "outputs must not alias inputs or each other");
    if (!(dx > 0.0) || !(dy > 0.0)) return fail(DCLAW_EINVAL, "dx and dy must be > 0");
    const int nstrips = (mx + STRIP_W - 1) / STRIP_W;
    const int seg_rows = g_seg_rows;
    const int nseg = (my + seg_rows - 1) / seg_rows;
    if ((long)nstrips * nseg > 0x7fffffffL) return fail(DCLAW_EINVAL, "patch too large for one launch");
    const int b0 = bc ? bc[0] : 0, b1 = bc ? bc[1] : 0, b2 = bc ? bc[2] : 0, b3 = bc ? bc[3] : 0;
    hipLaunchKernelGGL(k_stepgrid_tl, dim3(nstrips * nseg), dim3(64), 0, (hipStream_t)stream, q_in,
                       dq_in, q_out, dq_out, aux, maux, mx, my, dx, dy, dt_dev, params_dev,
                       dparams_dev, cfl_dev, seg_rows, nstrips, b0, b1, b2, b3);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

size_t dclaw_adjoint_workspace_doubles(int mx, int my)
{
    if (mx < 1 || my < 1) return 0;
    // the stage planes, the parameter-gradient partial sums, then the tile
    // marks, the active-tile list, its counter and the per-block counts (ints)
    const size_t ntiles = (size_t)((mx + 4 + ADJ_TW - 1) / ADJ_TW) *
                          (size_t)((my + 4 + ADJ_TH - 1) / ADJ_TH);
    const size_t ints = 2 * ntiles + 1 + (ntiles + 63) / 64;
    return (size_t)dclaw_adj::W_FIELDS * dclaw_adj::plane_size(mx, my) +
           (size_t)dclaw_adj::MAXP * ADJ_GP_PARTS + (ints + 1) / 2;
}

int dclaw_adjoint_step_dev(const double* q_in, const double* aux, int maux, int mx, int my,
                           double dx, double dy, const int* bc, const double* dt_dev,
                           const dclaw_params* params_dev, const dclaw_params* dparams_dev, int np,
                           const double* lam_out, double* lam_in, double* q_out, double* grad_dev,
                           double* work, size_t work_doubles, void* stream)
{
    if (int rcb = check_bc(bc)) return rcb;
    if (int rc = check_patch(mx, my, maux)) return rc;
    if (mx < 2 || my < 2) return fail(DCLAW_EINVAL, "the adjoint step needs mx, my >= 2");
    if (!q_in || !aux || !dt_dev || !params_dev || !lam_out || !lam_in || !work)
        return fail(DCLAW_EINVAL, "null pointer");
    if (np < 0 || np > dclaw_adj::MAXP) return fail(DCLAW_EINVAL, "np must be in [0, 4]");
    if (np > 0 && (!dparams_dev || !grad_dev))
        return fail(DCLAW_EINVAL, "parameter directions need dparams and grad");
    if ((q_out && (q_out == q_in || q_out == lam_out || q_out == lam_in)) || lam_in == q_in ||
        lam_in == lam_out)
        return fail(DCLAW_EINVAL, "q_out and lam_in must not alias q_in, lam_out or each other");
    if (!(dx > 0.0) || !(dy > 0.0)) return fail(DCLAW_EINVAL, "dx and dy must be > 0");
    if (work_doubles < dclaw_adjoint_workspace_doubles(mx, my))
        return fail(DCLAW_EINVAL, "adjoint workspace too small");
    dclaw_adj::AdjArgs A;
    A.q_in = q_in;
    A.aux = aux;
    A.maux = maux;
    A.mx = mx;
    A.my = my;
    A.dx = dx;
    A.dy = dy;
    A.dt = 0.0;
    for (int k = 0; k < 4; k++) A.bc[k] = bc ? bc[k] : 0;
    std::memset(&A.P, 0, sizeof(A.P));
    A.dP = dparams_dev;
    A.np = np;
    A.lam_out = lam_out;
    A.lam_in = lam_in;
    A.q_out = q_out;
    A.W = work;
    A.N = dclaw_adj::plane_size(mx, my);
    A.ld = dclaw_adj::plane_ld(mx);
    A.Pd = params_dev;
    A.dt_dev = dt_dev;
    const size_t nb = (A.N + ADJ_BLOCK - 1) / ADJ_BLOCK;
    if (nb > 0x7fffffffUL) return fail(DCLAW_EINVAL, "patch too large for one launch");
    const AdjTiles T = adj_tiles(work, mx, my, A.N);
    const int ntiles = T.ntx * T.nty;
    hipStream_t st = (hipStream_t)stream;
    const dim3 tg(adj_grid(ntiles)), tb(64);
    hipLaunchKernelGGL(k_zero_ints, dim3((ntiles + 255) / 256), dim3(256), 0, st, T.flags,
                       ntiles);
    hipLaunchKernelGGL(k_adjoint_dense, dim3((unsigned)nb), dim3(ADJ_BLOCK), 0, st, A, T);
    if (A.bc[0] != 0)
        hipLaunchKernelGGL(k_adjoint_border, dim3((ntiles + 255) / 256), dim3(256), 0, st, T,
                           mx, my);
    hipLaunchKernelGGL(k_adjoint_clear_stale, tg, tb, 0, st, A, T);
    const int nblk = (ntiles + 63) / 64;
    hipLaunchKernelGGL(k_adjoint_count_marks, dim3(nblk), dim3(64), 0, st, T, ntiles);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, T.bcount, nblk, T.count);
    hipLaunchKernelGGL(k_adjoint_list, dim3(nblk), dim3(64), 0, st, T, ntiles);
    hipLaunchKernelGGL(k_adjoint<ADJ_REC_PREPARE>, tg, tb, 0, st, A, T);
    hipLaunchKernelGGL(k_adjoint<ADJ_REC_TAUDIR>, tg, tb, 0, st, A, T);
    hipLaunchKernelGGL((k_adjointw<ADJ_REC_WAVES, 2>), tg, tb, 0, st, A, T);
    hipLaunchKernelGGL(k_adjoint<ADJ_REC_FLUX>, tg, tb, 0, st, A, T);
    hipLaunchKernelGGL(k_adjoint<ADJ_REC_SPLIT>, tg, tb, 0, st, A, T);
    hipLaunchKernelGGL((k_adjointw<ADJ_UPD_SRC2, 2>), tg, tb, 0, st, A, T);
    hipLaunchKernelGGL(k_adjoint<ADJ_REV_SRC2_SEEDED>, tg, tb, 0, st, A, T);
    hipLaunchKernelGGL((k_adjointw<ADJ_REV_SPLIT, 2>), tg, tb, 0, st, A, T);
    hipLaunchKernelGGL((k_adjointw<ADJ_REV_FLUX, 2>), tg, tb, 0, st, A, T);
    hipLaunchKernelGGL(k_adjoint<ADJ_REV_WAVES>, tg, tb, 0, st, A, T);
    hipLaunchKernelGGL((k_adjointw<ADJ_REV_TAUDIR, 2>), tg, tb, 0, st, A, T);
    hipLaunchKernelGGL((k_adjointw<ADJ_REV_PREPARE, 2>), tg, tb, 0, st, A, T);
    hipLaunchKernelGGL(k_adjoint<ADJ_REV_BC>, tg, tb, 0, st, A, T);
    if (np > 0) {
        double* part = work + (size_t)dclaw_adj::W_FIELDS * A.N;
        hipLaunchKernelGGL(k_adjoint_gp, dim3(ADJ_GP_PARTS, np), dim3(256), 0, st, work, A.N,
                           part);
        hipLaunchKernelGGL(k_adjoint_gp_final, dim3(np), dim3(ADJ_GP_PARTS), 0, st, part,
                           grad_dev);
    }
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

static long long adj_batch_tiles(const dclaw_adj_patch* ph, int npatch, long long* cells)
{
    long long t = 0, c = 0;
    for (int k = 0; k < npatch; k++) {
        const int mx = ph[k].mx, my = ph[k].my;
        c += (long long)dclaw_adj::plane_size(mx, my);
        t += (long long)((mx + 4 + ADJ_TW - 1) / ADJ_TW) * ((my + 4 + ADJ_TH - 1) / ADJ_TH);
    }
    if (cells) *cells = c;
    return t;
}

size_t dclaw_adjoint_batch_doubles(const dclaw_adj_patch* patches_host, int npatch)
{
    if (!patches_host || npatch < 1) return 0;
    const long long t = adj_batch_tiles(patches_host, npatch, nullptr);
    // cell and tile prefixes, the global counter, then the global list, the
    // per-block counts and the list's patch indices (ints)
    return (size_t)(2 * (npatch + 1) + 1) + (size_t)((2 * t + (t + 63) / 64 + 1) / 2);
}

int dclaw_adjoint_batched(const dclaw_adj_patch* patches_dev, const dclaw_adj_patch* patches_host,
                          int npatch, const int* bc, const double* dt_dev,
                          const dclaw_params* params_dev, const dclaw_params* dparams_dev, int np,
                          double* batch_work, size_t batch_doubles, void* stream)
{
    if (int rcb = check_bc(bc)) return rcb;
    if (!patches_dev || !patches_host || !dt_dev || !params_dev || !batch_work)
        return fail(DCLAW_EINVAL, "null pointer");
    if (npatch < 1 || npatch > 65535) return fail(DCLAW_EINVAL, "npatch must be in [1, 65535]");
    if (np < 0 || np > dclaw_adj::MAXP) return fail(DCLAW_EINVAL, "np must be in [0, 4]");
    if (np > 0 && !dparams_dev) return fail(DCLAW_EINVAL, "parameter directions need dparams");
    for (int k = 0; k < npatch; k++) {
        const dclaw_adj_patch& p = patches_host[k];
        if (int rc = check_patch(p.mx, p.my, p.maux)) return rc;
        if (p.mx < 2 || p.my < 2) return fail(DCLAW_EINVAL, "the adjoint step needs mx, my >= 2");
        if (!p.q_in || !p.aux || !p.lam_out || !p.lam_in || !p.work || (np > 0 && !p.grad))
            return fail(DCLAW_EINVAL, "null pointer in patch record");
        if ((p.q_out && (p.q_out == p.q_in || p.q_out == p.lam_out || p.q_out == p.lam_in)) ||
            p.lam_in == p.q_in || p.lam_in == p.lam_out)
            return fail(DCLAW_EINVAL, "q_out and lam_in must not alias q_in, lam_out or each other");
        if (!(p.dx > 0.0) || !(p.dy > 0.0)) return fail(DCLAW_EINVAL, "dx and dy must be > 0");
        if (p.member < 0) return fail(DCLAW_EINVAL, "negative member index");
    }
    if (batch_doubles < dclaw_adjoint_batch_doubles(patches_host, npatch))
        return fail(DCLAW_EINVAL, "adjoint batch workspace too small");
    long long ncells = 0;
    const long long ntiles = adj_batch_tiles(patches_host, npatch, &ncells);
    if (ntiles > 0x7fffffffLL || (ncells + ADJ_BLOCK - 1) / ADJ_BLOCK > 0x7fffffffLL)
        return fail(DCLAW_EINVAL, "batch too large for one launch");
    AdjBatch B;
    B.pt = patches_dev;
    B.npatch = npatch;
    for (int k = 0; k < 4; k++) B.bc[k] = bc ? bc[k] : 0;
    B.params = params_dev;
    B.dparams = np > 0 ? dparams_dev : nullptr;
    B.np = np;
    B.dt = dt_dev;
    B.cell_prefix = (long long*)batch_work;
    B.tile_prefix = B.cell_prefix + (npatch + 1);
    B.gcount = (int*)(B.tile_prefix + (npatch + 1));   // [0] active tiles, [1] see k_adjb_prefix
    const long long nblk = (ntiles + 63) / 64;
    B.glist = (int*)(batch_work + 2 * (npatch + 1) + 1);
    B.bcount = B.glist + ntiles;
    B.gmem = B.bcount + nblk;
    hipStream_t st = (hipStream_t)stream;
    const dim3 tg(adj_grid(ntiles)), tb(64);
    hipLaunchKernelGGL(k_adjb_prefix, dim3(1), dim3(1024), 0, st, B);
    hipLaunchKernelGGL(k_adjb_zero_flags, dim3(64, npatch), dim3(256), 0, st, B);
    hipLaunchKernelGGL(k_adjb_dense, dim3((unsigned)((ncells + ADJ_BLOCK - 1) / ADJ_BLOCK)),
                       dim3(ADJ_BLOCK), 0, st, B, ncells);
    if (B.bc[0] != 0)
        hipLaunchKernelGGL(k_adjb_border, dim3(64, npatch), dim3(256), 0, st, B);
    hipLaunchKernelGGL(k_adjb_clear_stale, dim3(32, npatch), tb, 0, st, B);
    hipLaunchKernelGGL(k_adjb_count_zero, dim3((npatch + 255) / 256), dim3(256), 0, st, B);
    hipLaunchKernelGGL(k_adjb_count_marks, dim3((unsigned)nblk), tb, 0, st, B, ntiles);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, B.bcount, (int)nblk, B.gcount);
    hipLaunchKernelGGL(k_adjb_list, dim3((unsigned)nblk), tb, 0, st, B, ntiles);
    hipLaunchKernelGGL(k_adjb<ADJ_REC_PREPARE>, tg, tb, 0, st, B);
    hipLaunchKernelGGL(k_adjb<ADJ_REC_TAUDIR>, tg, tb, 0, st, B);
    hipLaunchKernelGGL((k_adjbw<ADJ_REC_WAVES, 2>), tg, tb, 0, st, B);
    hipLaunchKernelGGL(k_adjb<ADJ_REC_FLUX>, tg, tb, 0, st, B);
    hipLaunchKernelGGL(k_adjb<ADJ_REC_SPLIT>, tg, tb, 0, st, B);
    hipLaunchKernelGGL((k_adjbw<ADJ_UPD_SRC2, 2>), tg, tb, 0, st, B);
    hipLaunchKernelGGL(k_adjb<ADJ_REV_SRC2_SEEDED>, tg, tb, 0, st, B);
    hipLaunchKernelGGL((k_adjbw<ADJ_REV_SPLIT, 2>), tg, tb, 0, st, B);
    hipLaunchKernelGGL((k_adjbw<ADJ_REV_FLUX, 2>), tg, tb, 0, st, B);
    hipLaunchKernelGGL(k_adjb<ADJ_REV_WAVES>, tg, tb, 0, st, B);
    hipLaunchKernelGGL((k_adjbw<ADJ_REV_TAUDIR, 2>), tg, tb, 0, st, B);
    hipLaunchKernelGGL((k_adjbw<ADJ_REV_PREPARE, 2>), tg, tb, 0, st, B);
    hipLaunchKernelGGL(k_adjb<ADJ_REV_BC>, tg, tb, 0, st, B);
    if (np > 0) {
        hipLaunchKernelGGL(k_adjb_gp, dim3(ADJ_GP_PARTS, np, npatch), dim3(256), 0, st, B);
        hipLaunchKernelGGL(k_adjb_gp_final, dim3(np, npatch), dim3(ADJ_GP_PARTS), 0, st, B);
    }
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_gauges_adj_dev(const double* q, const double* aux, int maux, int mx, int my,
                         double xlower, double ylower, double dx, double dy, double dry_tolerance,
                         const double* xy_dev, int ngauges, const double* lam_gauges_dev,
                         double* lam_q_dev, void* stream)
{
    if (int rc = check_patch(mx, my, maux)) return rc;
    if (!q || !aux || !xy_dev || !lam_gauges_dev || !lam_q_dev || ngauges < 1)
        return fail(DCLAW_EINVAL, "bad gauge args");
    if (!(dx > 0.0) || !(dy > 0.0)) return fail(DCLAW_EINVAL, "dx and dy must be > 0");
    hipLaunchKernelGGL(k_gauges_adj, dim3(1), dim3(64), 0, (hipStream_t)stream, q, aux, maux, mx,
                       my, xlower, ylower, dx, dy, dry_tolerance, xy_dev, ngauges, lam_gauges_dev,
                       lam_q_dev);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_tile_count(int mx, int my)
{
    if (mx < 1 || my < 1) return fail(DCLAW_EINVAL, "mx and my must be >= 1");
    const long n = tiles_at(mx, my, g_seg_rows.load());
    if (n > 0x7fffffff) return fail(DCLAW_EINVAL, "too many tiles");
    return (int)n;
}

int dclaw_split_resident_size(int mx, int my)
{
    const int n = dclaw_tile_count(mx, my);
    if (n < 0) return n;
    if (split_resident_ints(n) > 0x7fffffff) return fail(DCLAW_EINVAL, "too many tiles");
    return (int)split_resident_ints(n);
}

int dclaw_split_workspace_size(int mx, int my)
{
    const int n = dclaw_tile_count(mx, my);
    if (n < 0) return n;
    if (3L * n + 4 > 0x7fffffff) return fail(DCLAW_EINVAL, "too many tiles");
    return 3 * n + 4;
}

int dclaw_set_split_mode(int mode)
{
    if (mode < 0 || mode > 3) return fail(DCLAW_EINVAL, "split mode must be 0, 1, 2 or 3");
    return g_split_mode.exchange(mode);
}

int dclaw_set_dry_rows(int rows)
{
    if (rows < 1 || rows > 65536) return fail(DCLAW_EINVAL, "dry rows must be in [1, 65536]");
    return g_dry_rows.exchange(rows);
}

int dclaw_stepgrid_split_dev(const double* q_in, double* q_out, const double* aux, int maux,
                             int mx, int my, double dx, double dy, const int* bc,
                             const double* dt_dev, const dclaw_params* params_dev,
                             double* cfl_dev, int* tile_flags_dev, int n_flags, void* stream)
{
    if (!tile_flags_dev) return fail(DCLAW_EINVAL, "null tile_flags");
    return launch_stepgrid(global_tuning(), q_in, q_out, aux, nullptr, maux, mx, my, dx, dy,
                           dt_dev, params_dev, cfl_dev, bc, (hipStream_t)stream, tile_flags_dev,
                           n_flags);
}

}  // extern "C"

// Tile-flag workspace of dclaw_stepgrid_split_static_dev: a fixed array in
// the library's own device memory instead of a caller workspace, for callers
// that cannot allocate (graph capture, serving loops).  Each GPU the library
// runs on holds its own copy until the process ends; calls that overlap on
// one GPU share it and race.  The host never names the array in a kernel
// argument: it resolves the address once per device.
constexpr int STATIC_FLAG_INTS = 1 << 18;    // 1 MiB
__device__ __align__(256) int scratch[STATIC_FLAG_INTS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
int* g_scratch_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on the current device.
int static_scratch(int** out)
{
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(DCLAW_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<int*>(p);
    }
    *out = g_scratch_ptr[dev];
    return DCLAW_OK;
}
}  // namespace

extern "C" {

int dclaw_stepgrid_split_static_dev(const double* q_in, double* q_out, const double* aux,
                                    int maux, int mx, int my, double dx, double dy,
                                    const int* bc, const double* dt_dev,
                                    const dclaw_params* params_dev, double* cfl_dev,
                                    void* stream)
{
    const Tuning tn = global_tuning();
    if (mx < 1 || my < 1) return fail(DCLAW_EINVAL, "mx and my must be >= 1");
    const long ntiles = tiles_at(mx, my, tn.seg_rows);
    if (3L * ntiles + 4 > STATIC_FLAG_INTS)
        return fail(DCLAW_EINVAL, "patch has more tiles than the static workspace holds ("
                                  "dclaw_split_workspace_size > " +
                                      std::to_string(STATIC_FLAG_INTS) +
                                      " ints): use dclaw_stepgrid_split_dev");
    int* flags = nullptr;
    if (int rc = static_scratch(&flags)) return rc;
    return launch_stepgrid(tn, q_in, q_out, aux, nullptr, maux, mx, my, dx, dy, dt_dev,
                           params_dev, cfl_dev, bc, (hipStream_t)stream, flags,
                           (int)(3L * ntiles + 4));
}

int dclaw_static_workspace_ints(void) { return STATIC_FLAG_INTS; }

}  // extern "C"

// A caller's own library context: the settings its launches read instead of
// the library-wide ones, and the resident split step's tile workspace,
// allocated on the device current at the handle's first step and grown when
// a patch (or a shorter segment) needs more tiles.  One handle serves one
// time loop at a time: its calls are issued in order by one thread, on one
// stream, as the resident bookkeeping assumes.
struct dclaw_handle_s {
    Tuning tn;
    int* ws = nullptr;
    long ws_ints = 0;
    int device = -1;
};

extern "C" {

int dclaw_create(dclaw_handle* out)
{
    if (!out) return fail(DCLAW_EINVAL, "null handle pointer");
    *out = nullptr;
    dclaw_handle h = new (std::nothrow) dclaw_handle_s;
    if (!h) return fail(DCLAW_ENOMEM, "out of host memory");
    h->tn = global_tuning();
    *out = h;
    return DCLAW_OK;
}

int dclaw_destroy(dclaw_handle h)
{
    if (!h) return DCLAW_OK;
    int rc = DCLAW_OK;
    if (h->ws) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != h->device) e = hipSetDevice(h->device);
        // hipFree waits for the device's work, the last step included
        if (e == hipSuccess) e = hipFree(h->ws);
        if (e != hipSuccess) rc = hip_fail(e, "freeing the handle's workspace");
        if (cur >= 0 && cur != h->device) (void)hipSetDevice(cur);
    }
    delete h;
    return rc;
}

int dclaw_handle_set_seg_rows(dclaw_handle h, int seg_rows)
{
    if (!h) return fail(DCLAW_EINVAL, "null handle");
    if (seg_rows < 1 || seg_rows > 65536) return fail(DCLAW_EINVAL, "seg_rows must be in [1, 65536]");
    const int prev = h->tn.seg_rows;
    h->tn.seg_rows = seg_rows;
    return prev;
}

int dclaw_handle_set_split_mode(dclaw_handle h, int mode)
{
    if (!h) return fail(DCLAW_EINVAL, "null handle");
    if (mode < 0 || mode > 3) return fail(DCLAW_EINVAL, "split mode must be 0, 1, 2 or 3");
    const int prev = h->tn.split_mode;
    h->tn.split_mode = mode;
    return prev;
}

int dclaw_handle_set_dry_rows(dclaw_handle h, int rows)
{
    if (!h) return fail(DCLAW_EINVAL, "null handle");
    if (rows < 1 || rows > 65536) return fail(DCLAW_EINVAL, "dry rows must be in [1, 65536]");
    const int prev = h->tn.dry_rows;
    h->tn.dry_rows = rows;
    return prev;
}

int dclaw_handle_tile_count(dclaw_handle h, int mx, int my)
{
    if (!h) return fail(DCLAW_EINVAL, "null handle");
    if (mx < 1 || my < 1) return fail(DCLAW_EINVAL, "mx and my must be >= 1");
    const long n = tiles_at(mx, my, h->tn.seg_rows);
    if (split_resident_ints(n) > 0x7fffffff) return fail(DCLAW_EINVAL, "too many tiles");
    return (int)n;
}

int dclaw_handle_state_written(dclaw_handle h, void* stream)
{
    if (!h) return fail(DCLAW_EINVAL, "null handle");
    if (!h->ws) return DCLAW_OK;
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != h->device) return fail(DCLAW_EINVAL, "the handle's workspace is on another device");
    hipLaunchKernelGGL(k_zero_ints, dim3((unsigned)((h->ws_ints + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, h->ws, (int)h->ws_ints);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_handle_stepgrid_dev(dclaw_handle h, const double* q_in, double* q_out,
                              const double* aux, int maux, int mx, int my, double dx, double dy,
                              const int* bc, const double* dt_dev,
                              const dclaw_params* params_dev, double* cfl_dev, void* stream)
{
    if (!h) return fail(DCLAW_EINVAL, "null handle");
    const int ntiles = dclaw_handle_tile_count(h, mx, my);
    if (ntiles < 0) return ntiles;
    const long need = split_resident_ints(ntiles);
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (h->ws && dev != h->device)
        return fail(DCLAW_EINVAL, "the handle's workspace is on another device");
    const hipStream_t st = (hipStream_t)stream;
    if (h->ws_ints < need) {
        if (h->ws) {
            HIP_TRY(hipFree(h->ws));
            h->ws = nullptr;
            h->ws_ints = 0;
        }
        HIP_TRY(hipMalloc(&h->ws, (size_t)need * sizeof(int)));
        h->ws_ints = need;
        h->device = dev;
        // zero bookkeeping: the first step runs every tile
        hipLaunchKernelGGL(k_zero_ints, dim3((unsigned)((need + 255) / 256)), dim3(256), 0, st,
                           h->ws, (int)need);
        HIP_TRY(hipGetLastError());
    }
    return launch_stepgrid(h->tn, q_in, q_out, aux, nullptr, maux, mx, my, dx, dy, dt_dev,
                           params_dev, cfl_dev, bc, st, h->ws, (int)need);
}

size_t dclaw_curvature_workspace_doubles(int mx, int my)
{
    if (mx < 1 || my < 1) return 0;
    // the running gz of every interior cell, then one tile flag per
    // (strip, row) -- the most tiles any segment height gives
    const size_t flags = (size_t)((mx + STRIP_W - 1) / STRIP_W) * (size_t)my;
    return (size_t)mx * (size_t)my + (flags + 1) / 2;
}

static int launch_curvature_src2(double* q, const double* aux, int maux, int mx, int my,
                                 double dx, double dy, const double* dt_dev,
                                 const dclaw_params* params_dev, double* g, hipStream_t st)
{
    const long n = (long)mx * my;
    hipLaunchKernelGGL(k_curv_gacc, dim3(grid_for(n)), dim3(256), 0, st, q, aux, maux, mx, my,
                       dx, dy, params_dev, g);
    hipLaunchKernelGGL(k_curv_prefix, dim3(1), dim3(64), 0, st, g, n, params_dev);
    hipLaunchKernelGGL(k_src2_curv, dim3(grid_for(n)), dim3(256), 0, st, q, aux, maux, mx, my,
                       dt_dev, params_dev, g);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

static int check_stage(int mx, int my, int maux, bool need_taudir);

// The curvature step; aux_out (the host drop-in's aux, or null) receives
// aux(4..6) as b4step2's calc_taudir leaves them.
static int launch_curvature_step(const double* q_in, double* q_out, const double* aux,
                                 double* aux_out, int maux, int mx, int my, double dx, double dy,
                                 const int* bc, const double* dt_dev,
                                 const dclaw_params* params_dev, double* cfl_dev, double* work,
                                 size_t work_doubles, hipStream_t st)
{
    if (int rcb = check_bc(bc)) return rcb;
    if (int rc = check_patch(mx, my, maux)) return rc;
    if (maux < 3) return fail(DCLAW_EINVAL, "curvature needs maux >= 3 (aux(3) = theta)");
    if (!q_in || !q_out || !aux || !dt_dev || !params_dev || !cfl_dev || !work)
        return fail(DCLAW_EINVAL, "null pointer");
    if (q_in == q_out) return fail(DCLAW_EINVAL, "q_out must not alias q_in");
    if (!(dx > 0.0) || !(dy > 0.0)) return fail(DCLAW_EINVAL, "dx and dy must be > 0");
    if (work_doubles < dclaw_curvature_workspace_doubles(mx, my))
        return fail(DCLAW_EINVAL, "curvature workspace too small");
    const int nstrips = (mx + STRIP_W - 1) / STRIP_W;
    const int seg_rows = g_seg_rows;
    const int nseg = (my + seg_rows - 1) / seg_rows;
    const int b0 = bc ? bc[0] : 0, b1 = bc ? bc[1] : 0, b2 = bc ? bc[2] : 0, b3 = bc ? bc[3] : 0;
    int* flags = (int*)(work + (size_t)mx * (size_t)my);
    // the step without src2 (its dry reset included), then src2 with each
    // cell's running gz
    hipLaunchKernelGGL(k_stepgrid<MODE_STEP_NOSRC2>, dim3(nstrips * nseg), dim3(64), 0, st, q_in,
                       q_out, aux, aux_out, maux, mx, my, dx, dy, dt_dev, params_dev, cfl_dev,
                       seg_rows, nstrips, nseg, b0, b1, b2, b3, flags, nullptr, nullptr);
    HIP_TRY(hipGetLastError());
    return launch_curvature_src2(q_out, aux, maux, mx, my, dx, dy, dt_dev, params_dev, work, st);
}

int dclaw_stepgrid_curvature_dev(const double* q_in, double* q_out, const double* aux, int maux,
                                 int mx, int my, double dx, double dy, const int* bc,
                                 const double* dt_dev, const dclaw_params* params_dev,
                                 double* cfl_dev, double* work, size_t work_doubles,
                                 void* stream)
{
    return launch_curvature_step(q_in, q_out, aux, nullptr, maux, mx, my, dx, dy, bc, dt_dev,
                                 params_dev, cfl_dev, work, work_doubles, (hipStream_t)stream);
}

int dclaw_src2_curvature_dev(double* q, const double* aux, int maux, int mx, int my, double dx,
                             double dy, const double* dt_dev, const dclaw_params* params_dev,
                             double* work, size_t work_doubles, void* stream)
{
    if (int rc = check_stage(mx, my, maux, false)) return rc;
    if (maux < 3) return fail(DCLAW_EINVAL, "curvature needs maux >= 3 (aux(3) = theta)");
    if (!q || !aux || !dt_dev || !params_dev || !work) return fail(DCLAW_EINVAL, "null pointer");
    if (!(dx > 0.0) || !(dy > 0.0)) return fail(DCLAW_EINVAL, "dx and dy must be > 0");
    if (work_doubles < (size_t)mx * (size_t)my)
        return fail(DCLAW_EINVAL, "curvature workspace too small");
    return launch_curvature_src2(q, aux, maux, mx, my, dx, dy, dt_dev, params_dev, work,
                                 (hipStream_t)stream);
}

int dclaw_worklist_size(const int* mx, const int* my, int npatch, int seg_rows)
{
    if (!mx || !my || npatch < 0 || seg_rows < 1) return fail(DCLAW_EINVAL, "bad work-list args");
    long n = 0;
    for (int p = 0; p < npatch; p++) {
        if (mx[p] < 1 || my[p] < 1) return fail(DCLAW_EINVAL, "mx and my must be >= 1");
        n += (long)((mx[p] + STRIP_W - 1) / STRIP_W) * ((my[p] + seg_rows - 1) / seg_rows);
    }
    if (n > 0x7fffffff) return fail(DCLAW_EINVAL, "work list too long");
    return (int)n;
}

int dclaw_worklist_build(const int* mx, const int* my, int npatch, int seg_rows,
                         int32_t* tiles)
{
    const int n = dclaw_worklist_size(mx, my, npatch, seg_rows);
    if (n < 0) return n;
    if (!tiles) return fail(DCLAW_EINVAL, "null pointer");
    // Largest patches first so the long tiles start early; within a patch,
    // segments of one strip are consecutive.
    std::vector<int> order(npatch);
    for (int p = 0; p < npatch; p++) order[p] = p;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (long)mx[a] * my[a] > (long)mx[b] * my[b];
    });
    int t = 0;
    for (int p : order) {
        const int ns = (mx[p] + STRIP_W - 1) / STRIP_W;
        const int ng = (my[p] + seg_rows - 1) / seg_rows;
        for (int s = 0; s < ns; s++)
            for (int g = 0; g < ng; g++) {
                tiles[4 * t + 0] = p;
                tiles[4 * t + 1] = s;
                tiles[4 * t + 2] = g;
                tiles[4 * t + 3] = 0;
                t++;
            }
    }
    return n;
}

int dclaw_step_batched_ws(const dclaw_patch* patches_dev, int npatch, const int32_t* tiles_dev,
                          int ntiles, int seg_rows, const int* bc, const double* dt_dev,
                          const dclaw_params* params_dev, double* cfl_dev, int* tile_flags_dev,
                          int n_ws, void* stream)
{
    if (!patches_dev || !tiles_dev || !dt_dev || !params_dev || !cfl_dev)
        return fail(DCLAW_EINVAL, "null pointer");
    if (npatch < 1 || ntiles < 1 || seg_rows < 1) return fail(DCLAW_EINVAL, "empty batch");
    if (tile_flags_dev && n_ws < ntiles)
        return fail(DCLAW_EINVAL, "the workspace holds fewer ints than the work list has tiles");
    if (int rcb = check_bc(bc)) return rcb;
    const int b0 = bc ? bc[0] : 0, b1 = bc ? bc[1] : 0, b2 = bc ? bc[2] : 0, b3 = bc ? bc[3] : 0;
    const hipStream_t st = (hipStream_t)stream;
    if (!tile_flags_dev) {
        hipLaunchKernelGGL(k_step_batched<MODE_STEP>, dim3(ntiles), dim3(64), 0, st, patches_dev,
                           tiles_dev, seg_rows, dt_dev, params_dev, cfl_dev, b0, b1, b2, b3,
                           nullptr, nullptr, nullptr);
        HIP_TRY(hipGetLastError());
        return DCLAW_OK;
    }
    const int* order = nullptr;
    const int* dry_age = nullptr;
    if ((long)n_ws >= 2L * ntiles + 4) {
        const bool resident = (long)n_ws >= batched_resident_ints(ntiles, npatch);
        if (resident) {
            int* gmeta = tile_flags_dev + 2 * ntiles + 4;
            int* pmeta = gmeta + BSKIP_GMETA;
            if (b0 && b1 && b2 && b3) {
                // resident workspace, boundary conditions on load: tiles that
                // stayed dry are not read or written
                int* age = pmeta + (size_t)BSKIP_PMETA * npatch;
                hipLaunchKernelGGL(k_bskip_meta, dim3(1), dim3(256), 0, st, patches_dev, npatch,
                                   gmeta, pmeta, seg_rows, ntiles);
                hipLaunchKernelGGL(k_bdry_age, dim3((ntiles + 255) / 256), dim3(256), 0, st,
                                   patches_dev, tiles_dev, tile_flags_dev, age, pmeta, ntiles,
                                   seg_rows, (STRIP_HALO + seg_rows - 1) / seg_rows);
                dry_age = age;
            } else {
                // stored ghosts: every tile runs, and the next resident call
                // starts over
                hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, st, gmeta, BSKIP_GMETA);
            }
        }
        // the tiles the previous call flagged first
        int* ord = tile_flags_dev + ntiles;
        int* counters = tile_flags_dev + 2 * ntiles + 2;
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, st, counters, 2);
        hipLaunchKernelGGL(k_order_tiles, dim3((ntiles + 255) / 256), dim3(256), 0, st,
                           tile_flags_dev, ntiles, ord, counters);
        order = ord;
    }
    hipLaunchKernelGGL(k_step_batched<MODE_STEP_NOSRC2>, dim3(ntiles), dim3(64), 0, st,
                       patches_dev, tiles_dev, seg_rows, dt_dev, params_dev, cfl_dev, b0, b1, b2,
                       b3, tile_flags_dev, order, dry_age);
    HIP_TRY(hipGetLastError());
    const long ntask = (long)ntiles * ((STRIP_W * seg_rows + 63) / 64);
    hipLaunchKernelGGL(k_src2_batched, dim3((int)std::min<long>(ntask, 16384)), dim3(64), 0, st,
                       patches_dev, tiles_dev, ntiles, seg_rows, tile_flags_dev, dt_dev,
                       params_dev);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_batched_resident_size(int ntiles, int npatch)
{
    if (ntiles < 1 || npatch < 1) return fail(DCLAW_EINVAL, "empty batch");
    const long n = batched_resident_ints(ntiles, npatch);
    if (n > 0x7fffffff) return fail(DCLAW_EINVAL, "too many tiles");
    return (int)n;
}

int dclaw_step_batched(const dclaw_patch* patches_dev, int npatch, const int32_t* tiles_dev,
                       int ntiles, int seg_rows, const int* bc, const double* dt_dev,
                       const dclaw_params* params_dev, double* cfl_dev, int* tile_flags_dev,
                       void* stream)
{
    return dclaw_step_batched_ws(patches_dev, npatch, tiles_dev, ntiles, seg_rows, bc, dt_dev,
                                 params_dev, cfl_dev, tile_flags_dev, ntiles, stream);
}

int dclaw_step_batched_tl(const dclaw_patch* patches_dev, const dclaw_patch* tangents_dev,
                          int npatch, const int32_t* tiles_dev, int ntiles, int seg_rows,
                          const int* bc, const double* dt_dev, const dclaw_params* params_dev,
                          const dclaw_params* dparams_dev, double* cfl_dev, void* stream)
{
    if (!patches_dev || !tangents_dev || !tiles_dev || !dt_dev || !params_dev || !dparams_dev ||
        !cfl_dev)
        return fail(DCLAW_EINVAL, "null pointer");
    if (npatch < 1 || ntiles < 1 || seg_rows < 1) return fail(DCLAW_EINVAL, "empty batch");
    if (int rcb = check_bc(bc)) return rcb;
    const int b0 = bc ? bc[0] : 0, b1 = bc ? bc[1] : 0, b2 = bc ? bc[2] : 0, b3 = bc ? bc[3] : 0;
    hipLaunchKernelGGL(k_step_batched_tl, dim3(ntiles), dim3(64), 0, (hipStream_t)stream,
                       patches_dev, tangents_dev, tiles_dev, seg_rows, dt_dev, params_dev,
                       dparams_dev, cfl_dev, b0, b1, b2, b3);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_dt_update(double* dt_dev, double* cfl_dev, int n, double cfl_desired, double dt_max,
                    void* stream)
{
    if (!dt_dev || !cfl_dev || n < 1) return fail(DCLAW_EINVAL, "bad dt_update args");
    hipLaunchKernelGGL(k_dt_update, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       dt_dev, cfl_dev, n, cfl_desired, dt_max);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

static int check_stage(int mx, int my, int maux, bool need_taudir)
{
    int rc = check_patch(mx, my, maux);
    if (rc) return rc;
    if (need_taudir && maux < 6) return fail(DCLAW_EINVAL, "maux must be >= 6 (aux 4..6 hold fsphi, taudir)");
    if ((double)(mx + 4) * (double)(my + 4) * 7.0 >= 2147483647.0)
        return fail(DCLAW_EINVAL, "patch too large for the per-stage entry points");
    return DCLAW_OK;
}

int dclaw_check4nans_dev(const double* q, int mx, int my, int* first_nan_dev, void* stream)
{
    if (int rc = check_stage(mx, my, 1, false)) return rc;
    if (!q || !first_nan_dev) return fail(DCLAW_EINVAL, "null pointer");
    const long n = 7L * (mx + 4) * (my + 4);
    hipLaunchKernelGGL(k_check4nans, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, q, n,
                       first_nan_dev);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_b4step2_dev(double* q, double* aux, int maux, int mx, int my, double dx, double dy,
                      const dclaw_params* params_dev, int* first_nan_dev, void* stream)
{
    if (int rc = check_stage(mx, my, maux, true)) return rc;
    if (!q || !aux || !params_dev) return fail(DCLAW_EINVAL, "null pointer");
    if (!(dx > 0.0) || !(dy > 0.0)) return fail(DCLAW_EINVAL, "dx and dy must be > 0");
    const hipStream_t st = (hipStream_t)stream;
    if (first_nan_dev)
        if (int rc = dclaw_check4nans_dev(q, mx, my, first_nan_dev, stream)) return rc;
    const long n = (long)(mx + 4) * (my + 4);
    hipLaunchKernelGGL(k_b4step2_cells, dim3(grid_for(n)), dim3(256), 0, st, q, mx, my,
                       params_dev);
    HIP_TRY(hipGetLastError());
    const long nt = (long)(mx + 2) * (my + 2);
    hipLaunchKernelGGL(k_calc_taudir, dim3(grid_for(nt)), dim3(256), 0, st, q, aux, maux, mx, my,
                       dx, dy, params_dev);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_step2_dev(const double* q, const double* aux, int maux, int mx, int my, double dx,
                    double dy, const double* dt_dev, const dclaw_params* params_dev, double* fm,
                    double* fp, double* gm, double* gp, double* cfl_dev, void* stream)
{
    if (int rc = check_stage(mx, my, maux, true)) return rc;
    if (!q || !aux || !dt_dev || !params_dev || !fm || !fp || !gm || !gp || !cfl_dev)
        return fail(DCLAW_EINVAL, "null pointer");
    if (!(dx > 0.0) || !(dy > 0.0)) return fail(DCLAW_EINVAL, "dx and dy must be > 0");
    const hipStream_t st = (hipStream_t)stream;
    const long nq = 7L * (mx + 4) * (my + 4);
    double* outs[4] = {fm, fp, gm, gp};
    for (double* a : outs) {
        hipLaunchKernelGGL(k_fill, dim3(grid_for(nq)), dim3(256), 0, st, a, nq, 0.0);
        HIP_TRY(hipGetLastError());
    }
    const int nstrips = (mx + STRIP_W - 1) / STRIP_W;
    const int seg_rows = g_seg_rows;
    const int nseg = (my + seg_rows - 1) / seg_rows;
    if (nseg > 65535) return fail(DCLAW_EINVAL, "my too large for one launch");
    hipLaunchKernelGGL(k_step2, dim3(nstrips, nseg), dim3(64), 0, st, q, aux, maux, mx, my, dx, dy,
                       dt_dev, params_dev, fm, fp, gm, gp, cfl_dev, seg_rows);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_src2_dev(double* q, const double* aux, int maux, int mx, int my, const double* dt_dev,
                   const dclaw_params* params_dev, int* first_nan_dev, void* stream)
{
    if (int rc = check_stage(mx, my, maux, false)) return rc;
    if (!q || !aux || !dt_dev || !params_dev) return fail(DCLAW_EINVAL, "null pointer");
    if (first_nan_dev)
        if (int rc = dclaw_check4nans_dev(q, mx, my, first_nan_dev, stream)) return rc;
    const long n = (long)mx * my;
    hipLaunchKernelGGL(k_src2, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, q, aux, maux,
                       mx, my, dt_dev, params_dev);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_dt_update_out(double* dt_dev, double* cfl_dev, double* t_dev, const double* tout_dev,
                        int* hit_dev, int n, double cfl_desired, double dt_max, void* stream)
{
    if (!dt_dev || !cfl_dev || !t_dev || !tout_dev || !hit_dev || n < 1)
        return fail(DCLAW_EINVAL, "bad dt_update_out args");
    hipLaunchKernelGGL(k_dt_update_out, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       dt_dev, cfl_dev, t_dev, tout_dev, hit_dev, n, cfl_desired, dt_max);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_gauges_dev(const double* q, const double* aux, int maux, int mx, int my, double xlower,
                     double ylower, double dx, double dy, double dry_tolerance,
                     const double* xy_dev, int ngauges, double* out_dev, void* stream)
{
    if (int rc = check_patch(mx, my, maux)) return rc;
    if (!q || !aux || !xy_dev || !out_dev || ngauges < 1) return fail(DCLAW_EINVAL, "bad gauge args");
    if (!(dx > 0.0) || !(dy > 0.0)) return fail(DCLAW_EINVAL, "dx and dy must be > 0");
    hipLaunchKernelGGL(k_gauges, dim3((ngauges + 63) / 64), dim3(64), 0, (hipStream_t)stream, q,
                       aux, maux, mx, my, xlower, ylower, dx, dy, dry_tolerance, xy_dev, ngauges,
                       out_dev, nullptr, nullptr);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_gauges_tl_dev(const double* q, const double* dq, const double* aux, int maux, int mx,
                        int my, double xlower, double ylower, double dx, double dy,
                        double dry_tolerance, const double* xy_dev, int ngauges, double* out_dev,
                        double* dout_dev, void* stream)
{
    if (int rc = check_patch(mx, my, maux)) return rc;
    if (!q || !dq || !aux || !xy_dev || !out_dev || !dout_dev || ngauges < 1)
        return fail(DCLAW_EINVAL, "bad gauge args");
    if (!(dx > 0.0) || !(dy > 0.0)) return fail(DCLAW_EINVAL, "dx and dy must be > 0");
    hipLaunchKernelGGL(k_gauges, dim3((ngauges + 63) / 64), dim3(64), 0, (hipStream_t)stream, q,
                       aux, maux, mx, my, xlower, ylower, dx, dy, dry_tolerance, xy_dev, ngauges,
                       out_dev, dq, dout_dev);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_amr_average_dev(double* qc, const double* auxc, int mauxc, int mitc, int mjtc,
                          const double* qf, const double* auxf, int mauxf, int mitf, int mjtf,
                          int ci0, int cj0, int ni, int nj, int fi0, int fj0, int rx, int ry,
                          double dry_tolerance, void* stream)
{
    if (!qc || !auxc || !qf || !auxf) return fail(DCLAW_EINVAL, "null pointer");
    if (mauxc < 1 || mauxf < 1 || rx < 1 || ry < 1 || ni < 0 || nj < 0)
        return fail(DCLAW_EINVAL, "bad AMR average sizes");
    if (ci0 < 0 || cj0 < 0 || (long)ci0 + ni > mitc || (long)cj0 + nj > mjtc ||
        fi0 < 0 || fj0 < 0 || (long)fi0 + (long)ni * rx > mitf || (long)fj0 + (long)nj * ry > mjtf)
        return fail(DCLAW_EINVAL, "AMR average block outside the patch arrays");
    if (!(dry_tolerance >= 0.0)) return fail(DCLAW_EINVAL, "dry_tolerance must be >= 0");
    const long n = (long)ni * nj;
    if (n == 0) return DCLAW_OK;
    const int blocks = (int)std::min<long>((n + 255) / 256, 65535);
    hipLaunchKernelGGL(k_amr_average, dim3(blocks), dim3(256), 0, (hipStream_t)stream, qc, auxc,
                       mauxc, mitc, qf, auxf, mauxf, mitf, ci0, cj0, ni, nj, fi0, fj0, rx, ry,
                       dry_tolerance);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_amr_interp_dev(const double* valc, const double* auxc, int mauxc, int mic, int mjc,
                         double* val, const double* auxf, int mauxf, int mitot, int mjtot,
                         const unsigned char* needs_set, int rx, int ry, double dry_tolerance,
                         double sea_level, double rho_f, double grav, void* stream)
{
    if (!valc || !auxc || !val || !auxf || !needs_set) return fail(DCLAW_EINVAL, "null pointer");
    if (mauxc < 1 || mauxf < 1 || rx < 1 || ry < 1 || mic < 3 || mjc < 3)
        return fail(DCLAW_EINVAL, "bad AMR interpolation sizes");
    if ((long)(mic - 2) * rx + 4 > mitot || (long)(mjc - 2) * ry + 4 > mjtot)
        return fail(DCLAW_EINVAL, "fine patch smaller than the coarse patch's interior blocks");
    if (!(dry_tolerance >= 0.0)) return fail(DCLAW_EINVAL, "dry_tolerance must be >= 0");
    const long n = (long)(mic - 2) * (mjc - 2);
    const int blocks = (int)std::min<long>((n + 255) / 256, 65535);
    hipLaunchKernelGGL(k_amr_interp, dim3(blocks), dim3(256), 0, (hipStream_t)stream, valc, auxc,
                       mauxc, mic, mjc, val, auxf, mauxf, mitot, needs_set, rx, ry, dry_tolerance,
                       sea_level, rho_f, grav);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

static int check_region(const dclaw_amr_region* r)
{
    if (!r || !r->q || !r->aux || !r->flag) return fail(DCLAW_EINVAL, "null region pointer");
    if (r->maux < 1 || r->mx < 1 || r->my < 1 || r->ihi < r->ilo || r->jhi < r->jlo)
        return fail(DCLAW_EINVAL, "bad AMR region sizes");
    if (r->r0 < 0 || r->c0 < 0 || (long)r->r0 + (r->ihi - r->ilo) >= r->mx ||
        (long)r->c0 + (r->jhi - r->jlo) >= r->my)
        return fail(DCLAW_EINVAL, "AMR region outside its arrays");
    return DCLAW_OK;
}

static int amr_blocks(long n) { return (int)std::max<long>(1, std::min<long>((n + 255) / 256, 65535)); }

int dclaw_amr_region_init_dev(const dclaw_amr_region* r, int iregsz, int jregsz, void* stream)
{
    if (int rc = check_region(r)) return rc;
    if (iregsz < 1 || jregsz < 1) return fail(DCLAW_EINVAL, "bad domain size");
    const long n = (long)(r->ihi - r->ilo + 1) * (r->jhi - r->jlo + 1);
    hipLaunchKernelGGL(k_amr_region_init, dim3(amr_blocks(n)), dim3(256), 0,
                       (hipStream_t)stream, *r, iregsz, jregsz);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_amr_copy_dev(const dclaw_amr_region* r, const double* src_q, const double* src_q_old,
                       double w_old, const double* src_aux, int src_maux, int src_ilo,
                       int src_jlo, int src_mx, int src_my, int copy_aux, void* stream)
{
    if (int rc = check_region(r)) return rc;
    if (!src_q || (copy_aux && !src_aux)) return fail(DCLAW_EINVAL, "null source pointer");
    if (src_mx < 1 || src_my < 1 || (copy_aux && src_maux < 1))
        return fail(DCLAW_EINVAL, "bad source patch sizes");
    if (!(w_old >= 0.0 && w_old <= 1.0)) return fail(DCLAW_EINVAL, "w_old must be in [0, 1]");
    if (src_q == r->q && (src_q_old || copy_aux))
        return fail(DCLAW_EINVAL, "a patch marking its own cells copies nothing");
    const int ia = std::max(r->ilo, src_ilo), ib = std::min(r->ihi, src_ilo + src_mx - 1);
    const int ja = std::max(r->jlo, src_jlo), jb = std::min(r->jhi, src_jlo + src_my - 1);
    if (ia > ib || ja > jb) return DCLAW_OK;
    const long n = (long)(ib - ia + 1) * (jb - ja + 1);
    hipLaunchKernelGGL(k_amr_copy, dim3(amr_blocks(n)), dim3(256), 0, (hipStream_t)stream, *r,
                       src_q, src_q_old, w_old, src_aux, src_maux, src_ilo, src_jlo, src_mx,
                       copy_aux, ia, ib, ja, jb);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_amr_unset_box_dev(const dclaw_amr_region* r, int* box_dev, void* stream)
{
    if (int rc = check_region(r)) return rc;
    if (!box_dev) return fail(DCLAW_EINVAL, "null box pointer");
    hipLaunchKernelGGL(k_amr_unset_box, dim3(1), dim3(1024), 0, (hipStream_t)stream, *r,
                       box_dev);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

void dclaw_amr_coarse_box(int isl, int isr, int jsb, int jst, int rx, int ry, int* box)
{
    if (box && rx > 0 && ry > 0) amr_coarse_box(isl, isr, jsb, jst, rx, ry, box);
}

int dclaw_amr_filpatch_dev(const dclaw_amr_region* r, const double* valc, const double* auxc,
                           int mauxc, int iplo, int iphi, int jplo, int jphi, int rx, int ry,
                           double xlower, double ylower, double dx_fine, double dy_fine,
                           double dx_coarse, double dy_coarse, double sea_level,
                           double dry_tolerance, double rho_f, double grav, void* stream)
{
    if (int rc = check_region(r)) return rc;
    if (!valc || !auxc) return fail(DCLAW_EINVAL, "null coarse pointer");
    if (mauxc < 1 || rx < 1 || ry < 1 || iphi - iplo < 2 || jphi - jplo < 2)
        return fail(DCLAW_EINVAL, "bad coarse rectangle");
    if (!(dx_fine > 0.0 && dy_fine > 0.0 && dx_coarse > 0.0 && dy_coarse > 0.0))
        return fail(DCLAW_EINVAL, "cell sizes must be > 0");
    if (!(dry_tolerance >= 0.0)) return fail(DCLAW_EINVAL, "dry_tolerance must be >= 0");
    AmrFill G;
    G.ilo = r->ilo; G.ihi = r->ihi; G.jlo = r->jlo; G.jhi = r->jhi;
    G.r0 = r->r0; G.c0 = r->c0; G.mx = r->mx; G.maux = r->maux;
    G.iplo = iplo; G.jplo = jplo; G.mxc = iphi - iplo + 1; G.myc = jphi - jplo + 1;
    G.mauxc = mauxc; G.rx = rx; G.ry = ry;
    G.xlower = xlower; G.ylower = ylower; G.dxf = dx_fine; G.dyf = dy_fine;
    G.dxc = dx_coarse; G.dyc = dy_coarse;
    G.sea_level = sea_level; G.drytol = dry_tolerance; G.rho_f = rho_f; G.grav = grav;
    const long n = (long)(G.mxc - 2) * (G.myc - 2);
    hipLaunchKernelGGL(k_amr_filpatch, dim3(amr_blocks(n)), dim3(256), 0, (hipStream_t)stream,
                       *r, valc, auxc, G);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_amr_region_bc_dev(const dclaw_amr_region* r, int iregsz, int jregsz, const int* bc,
                            void* stream)
{
    if (int rc = check_region(r)) return rc;
    if (int rcb = check_bc(bc)) return rcb;
    if (!bc) return fail(DCLAW_EINVAL, "null bc");
    if (iregsz < 1 || jregsz < 1) return fail(DCLAW_EINVAL, "bad domain size");
    // every source cell the rules name must be a domain cell of the region
    const int xs[2] = {r->ilo, r->ihi}, ys[2] = {r->jlo, r->jhi};
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            int si, sj;
            double fu, fv;
            amr_domain_source(bc, iregsz, jregsz, xs[a], ys[b], si, sj, fu, fv);
            if (si < r->ilo || si > r->ihi || sj < r->jlo || sj > r->jhi ||
                !amr_in_domain(si, sj, iregsz, jregsz))
                return fail(DCLAW_EINVAL, "AMR region: boundary source cells outside the region");
        }
    const long n = (long)(r->ihi - r->ilo + 1) * (r->jhi - r->jlo + 1);
    hipLaunchKernelGGL(k_amr_region_bc, dim3(amr_blocks(n)), dim3(256), 0, (hipStream_t)stream,
                       *r, iregsz, jregsz, bc[0], bc[1], bc[2], bc[3]);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_amr_setaux_dev(double* aux, int maux, int mx, int my, int ilo, int jlo,
                         const unsigned char* bed_set, const double* bed, int iregsz,
                         int jregsz, double phi, double theta, void* stream)
{
    if (!aux || !bed_set || !bed) return fail(DCLAW_EINVAL, "null pointer");
    if (maux < 3 || mx < 1 || my < 1) return fail(DCLAW_EINVAL, "setaux needs maux >= 3");
    if (iregsz < 1 || jregsz < 1) return fail(DCLAW_EINVAL, "bad domain size");
    if (ilo + mx - 1 < 0 || ilo > iregsz - 1 || jlo + my - 1 < 0 || jlo > jregsz - 1)
        return fail(DCLAW_EINVAL, "patch outside the domain");
    const long n = (long)(mx + 4) * (my + 4);
    hipLaunchKernelGGL(k_amr_setaux_bed, dim3(amr_blocks(n)), dim3(256), 0, (hipStream_t)stream,
                       aux, maux, mx, my, ilo, jlo, bed_set, bed, iregsz, jregsz);
    hipLaunchKernelGGL(k_amr_setaux_fields, dim3(amr_blocks(n)), dim3(256), 0,
                       (hipStream_t)stream, aux, maux, mx, my, ilo, jlo, iregsz, jregsz, phi,
                       theta);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

int dclaw_set_params(const dclaw_params* p)
{
    if (!p) return fail(DCLAW_EINVAL, "null params");
    int rc = validate_params(*p);
    if (rc) return rc;
    std::unique_lock<std::shared_mutex> lk(g_cfg.mu);
    if (!g_cfg.d_params) HIP_TRY(hipMalloc(&g_cfg.d_params, sizeof(dclaw_params)));
    HIP_TRY(hipMemcpy(g_cfg.d_params, p, sizeof(dclaw_params), hipMemcpyHostToDevice));
    g_cfg.params = *p;
    g_cfg.have_params = true;
    return DCLAW_OK;
}

// Host <-> device copies on the calling thread's stream.
static int h2d(HostCtx* c, double* d, const double* h, size_t n)
{
    HIP_TRY(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, c->st));
    return DCLAW_OK;
}
static int d2h(HostCtx* c, double* h, const double* d, size_t n)
{
    HIP_TRY(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
    return DCLAW_OK;
}

static int stepgrid_host(double* q, double* aux, int mbc, int mx, int my, int maux, double dx,
                         double dy, double dt, double* cflgrid, double* const* flux = nullptr)
{
    if (mbc != DCLAW_MBC) return fail(DCLAW_EINVAL, "mbc must be 2");
    int rc = check_patch(mx, my, maux);
    if (rc) return rc;
    if (maux < 6) return fail(DCLAW_EINVAL, "the host drop-in needs maux >= 6");
    std::shared_lock<std::shared_mutex> lk(g_cfg.mu);
    if (!g_cfg.have_params) return fail(DCLAW_EINVAL, "dclaw_set_params was not called");
    if (g_cfg.params.entrainment && maux < 7)
        return fail(DCLAW_EINVAL, "entrainment needs maux >= 7 (aux(7) holds h_e)");
    HostCtx* c;
    if ((rc = host_ctx(&c))) return rc;
    const size_t nq = (size_t)7 * (mx + 4) * (my + 4);
    const size_t na = (size_t)maux * (mx + 4) * (my + 4);
    if ((rc = grow(&c->d_qa, &c->cap_qa, nq))) return rc;
    if ((rc = grow(&c->d_qb, &c->cap_qb, nq))) return rc;
    if ((rc = grow(&c->d_aux, &c->cap_aux, na))) return rc;
    double scal[2] = {dt, 0.0};
    if ((rc = h2d(c, c->d_qa, q, nq)) || (rc = h2d(c, c->d_aux, aux, na)) ||
        (rc = h2d(c, c->d_scalars, scal, 2)))
        return rc;
    if (flux) {
        // fm, fp, gm, gp of the step: step2 on the b4step2 state of a copy
        // (the fused step below computes the same values internally)
        for (int k = 0; k < 4; k++)
            if ((rc = grow(&c->d_flux[k], &c->cap_flux[k], nq))) return rc;
        if ((rc = grow(&c->d_qc, &c->cap_qc, nq))) return rc;
        HIP_TRY(hipMemcpyAsync(c->d_qc, c->d_qa, nq * sizeof(double), hipMemcpyDeviceToDevice,
                               c->st));
        if ((rc = dclaw_b4step2_dev(c->d_qc, c->d_aux, maux, mx, my, dx, dy, g_cfg.d_params,
                                    nullptr, c->st)))
            return rc;
        if ((rc = dclaw_step2_dev(c->d_qc, c->d_aux, maux, mx, my, dx, dy, c->d_scalars,
                                  g_cfg.d_params, c->d_flux[0], c->d_flux[1], c->d_flux[2],
                                  c->d_flux[3], c->d_scalars + 1, c->st)))
            return rc;
    }
    if (g_cfg.params.curvature == 1) {
        const size_t nw = dclaw_curvature_workspace_doubles(mx, my);
        if ((rc = grow(&c->d_curv, &c->cap_curv, nw))) return rc;
        rc = launch_curvature_step(c->d_qa, c->d_qb, c->d_aux, c->d_aux, maux, mx, my, dx, dy,
                                   nullptr, c->d_scalars, g_cfg.d_params, c->d_scalars + 1,
                                   c->d_curv, nw, c->st);
    } else {
        rc = launch_stepgrid(global_tuning(), c->d_qa, c->d_qb, c->d_aux, c->d_aux, maux, mx,
                             my, dx, dy, c->d_scalars, g_cfg.d_params, c->d_scalars + 1, nullptr,
                             c->st);
    }
    if (rc) return rc;
    if ((rc = d2h(c, q, c->d_qb, nq)) || (rc = d2h(c, aux, c->d_aux, na)) ||
        (rc = d2h(c, scal, c->d_scalars, 2)))
        return rc;
    if (flux)
        for (int k = 0; k < 4; k++)
            if ((rc = d2h(c, flux[k], c->d_flux[k], nq))) return rc;
    HIP_TRY(hipStreamSynchronize(c->st));
    *cflgrid = scal[1];
    return DCLAW_OK;
}

// ---- Fortran-ABI per-stage drop-ins (host memory) ----------------------

// Copies q (and aux) of one patch to the calling thread's device buffers.
static int host_upload(HostCtx* c, const double* q, const double* aux, int mx, int my, int maux)
{
    int rc;
    const size_t nq = (size_t)7 * (mx + 4) * (my + 4);
    const size_t na = (size_t)maux * (mx + 4) * (my + 4);
    if ((rc = grow(&c->d_qa, &c->cap_qa, nq))) return rc;
    if ((rc = grow(&c->d_aux, &c->cap_aux, na))) return rc;
    if ((rc = h2d(c, c->d_qa, q, nq))) return rc;
    if (aux && (rc = h2d(c, c->d_aux, aux, na))) return rc;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, c->st, c->d_nan, 0x7fffffff);
    HIP_TRY(hipGetLastError());
    return DCLAW_OK;
}

// check4nans' stop (AMRClaw check4nans): the first NaN found by the scan.
static void host_nan_stop(HostCtx* c, const double* q, int mx, int my, double t, int icheck)
{
    int first = 0x7fffffff;
    if (hipMemcpyAsync(&first, c->d_nan, sizeof(int), hipMemcpyDeviceToHost, c->st) !=
            hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess)
        return;
    if (first == 0x7fffffff) return;
    const int m = first % 7, cell = first / 7;
    const int i = cell % (mx + 4) - 1, j = cell / (mx + 4) - 1;
    std::fprintf(stderr, "SOLUTION ERROR --- ABORTING CALCULATION\n");
    std::fprintf(stderr, "At icheck = %d\n", icheck);
    std::fprintf(stderr, "    mx,my,t: %d %d %.17g\n", mx, my, t);
    std::fprintf(stderr, "    m,i,j: %d %d %d\n", m + 1, i, j);
    std::fprintf(stderr, "    q(m,i,j) = %g\n", q[first]);
    std::abort();
}

static int check_host_stage(int mbc, int meqn, int mx, int my, int maux, bool need_taudir)
{
    if (mbc != DCLAW_MBC) return fail(DCLAW_EINVAL, "mbc must be 2");
    if (meqn != DCLAW_MEQN) return fail(DCLAW_EINVAL, "meqn must be 7");
    if (int rc = check_stage(mx, my, maux, need_taudir)) return rc;
    if (!g_cfg.have_params) return fail(DCLAW_EINVAL, "dclaw_set_params was not called");
    if (g_cfg.params.entrainment && maux < 7)
        return fail(DCLAW_EINVAL, "entrainment needs maux >= 7 (aux(7) holds h_e)");
    if (g_cfg.params.curvature && maux < 3)
        return fail(DCLAW_EINVAL, "curvature needs maux >= 3 (aux(3) = theta)");
    return DCLAW_OK;
}

static void host_stop_on_error(const char* who, int rc)
{
    if (rc == DCLAW_OK) return;
    std::fprintf(stderr, "%s: %s\n", who, g_err.c_str());
    std::abort();
}

static int b4step2_host(int mbc, int mx, int my, int meqn, double* q, double dx, double dy,
                        double t, int maux, double* aux)
{
    std::shared_lock<std::shared_mutex> lk(g_cfg.mu);
    if (int rc = check_host_stage(mbc, meqn, mx, my, maux, true)) return rc;
    HostCtx* c;
    if (int rc = host_ctx(&c)) return rc;
    if (int rc = host_upload(c, q, aux, mx, my, maux)) return rc;
    if (int rc = dclaw_b4step2_dev(c->d_qa, c->d_aux, maux, mx, my, dx, dy, g_cfg.d_params,
                                   c->d_nan, c->st))
        return rc;
    host_nan_stop(c, q, mx, my, t, 1);
    const size_t nq = (size_t)7 * (mx + 4) * (my + 4);
    const size_t na = (size_t)maux * (mx + 4) * (my + 4);
    if (int rc = d2h(c, q, c->d_qa, nq)) return rc;
    if (int rc = d2h(c, aux, c->d_aux, na)) return rc;
    HIP_TRY(hipStreamSynchronize(c->st));
    return DCLAW_OK;
}

static int step2_host(int meqn, int maux, int mbc, int mx, int my, const double* qold,
                      const double* aux, double dx, double dy, double dt, double* cflgrid,
                      double* fm, double* fp, double* gm, double* gp)
{
    std::shared_lock<std::shared_mutex> lk(g_cfg.mu);
    if (int rc = check_host_stage(mbc, meqn, mx, my, maux, true)) return rc;
    HostCtx* c;
    if (int rc = host_ctx(&c)) return rc;
    if (int rc = host_upload(c, qold, aux, mx, my, maux)) return rc;
    const size_t nq = (size_t)7 * (mx + 4) * (my + 4);
    for (int k = 0; k < 4; k++)
        if (int rc = grow(&c->d_flux[k], &c->cap_flux[k], nq)) return rc;
    double scal[2] = {dt, 0.0};
    if (int rc = h2d(c, c->d_scalars, scal, 2)) return rc;
    if (int rc = dclaw_step2_dev(c->d_qa, c->d_aux, maux, mx, my, dx, dy, c->d_scalars,
                                 g_cfg.d_params, c->d_flux[0], c->d_flux[1], c->d_flux[2],
                                 c->d_flux[3], c->d_scalars + 1, c->st))
        return rc;
    double* outs[4] = {fm, fp, gm, gp};
    for (int k = 0; k < 4; k++)
        if (int rc = d2h(c, outs[k], c->d_flux[k], nq)) return rc;
    if (int rc = d2h(c, scal, c->d_scalars, 2)) return rc;
    HIP_TRY(hipStreamSynchronize(c->st));
    *cflgrid = scal[1];
    return DCLAW_OK;
}

static int src2_host(int meqn, int mbc, int mx, int my, double* q, int maux, const double* aux,
                     double dx, double dy, double t, double dt)
{
    std::shared_lock<std::shared_mutex> lk(g_cfg.mu);
    if (int rc = check_host_stage(mbc, meqn, mx, my, maux, false)) return rc;
    HostCtx* c;
    if (int rc = host_ctx(&c)) return rc;
    // aux travels with q: with entrainment on src2 reads aux(7) = h_e
    // (src2.f90:253-263), and the device buffer may hold another patch's aux
    if (int rc = host_upload(c, q, aux, mx, my, maux)) return rc;
    double scal[2] = {dt, 0.0};
    if (int rc = h2d(c, c->d_scalars, scal, 2)) return rc;
    if (g_cfg.params.curvature == 1) {
        const size_t nw = (size_t)mx * (size_t)my;
        if (int rc = grow(&c->d_curv, &c->cap_curv, nw)) return rc;
        if (int rc = dclaw_check4nans_dev(c->d_qa, mx, my, c->d_nan, c->st)) return rc;
        if (int rc = dclaw_src2_curvature_dev(c->d_qa, c->d_aux, maux, mx, my, dx, dy,
                                              c->d_scalars, g_cfg.d_params, c->d_curv, nw,
                                              c->st))
            return rc;
    } else if (int rc = dclaw_src2_dev(c->d_qa, c->d_aux, maux, mx, my, c->d_scalars,
                                       g_cfg.d_params, c->d_nan, c->st)) {
        return rc;
    }
    host_nan_stop(c, q, mx, my, t, 2);
    const size_t nq = (size_t)7 * (mx + 4) * (my + 4);
    if (int rc = d2h(c, q, c->d_qa, nq)) return rc;
    HIP_TRY(hipStreamSynchronize(c->st));
    return DCLAW_OK;
}

void b4step2_(const int* mbc, const int* mx, const int* my, const int* meqn, double* q,
              const double* xlower, const double* ylower, const double* dx, const double* dy,
              const double* t, const double* dt, const int* maux, double* aux,
              const int* actualstep)
{
    (void)xlower; (void)ylower; (void)dt; (void)actualstep;
    host_stop_on_error("b4step2", b4step2_host(*mbc, *mx, *my, *meqn, q, *dx, *dy, *t, *maux, aux));
}

void step2_(const int* maxm, const int* meqn, const int* maux, const int* mbc, const int* mx,
            const int* my, const double* qold, const double* aux, const double* dx,
            const double* dy, const double* dt, double* cflgrid, double* fm, double* fp,
            double* gm, double* gp, void* rpn2, void* rpt2)
{
    (void)maxm; (void)rpn2; (void)rpt2;
    host_stop_on_error("step2", step2_host(*meqn, *maux, *mbc, *mx, *my, qold, aux, *dx, *dy, *dt,
                                           cflgrid, fm, fp, gm, gp));
}

void src2_(const int* meqn, const int* mbc, const int* mx, const int* my, const double* xlower,
           const double* ylower, const double* dx, const double* dy, double* q, const int* maux,
           const double* aux, const double* t, const double* dt)
{
    (void)xlower; (void)ylower;
    host_stop_on_error("src2", src2_host(*meqn, *mbc, *mx, *my, q, *maux, aux, *dx, *dy, *t, *dt));
}

void stepgrid_dclaw_(double* q, double* aux, const int* mbc, const int* mx, const int* my,
                     const int* maux, const double* dx, const double* dy, const double* dt,
                     double* cflgrid)
{
    const int rc = stepgrid_host(q, aux, *mbc, *mx, *my, *maux, *dx, *dy, *dt, cflgrid);
    if (rc != DCLAW_OK) {
        // The Fortran caller has no status argument; like the reference's
        // fatal paths (check4nans, set_dig) this stops the run.
        std::fprintf(stderr, "stepgrid_dclaw: %s\n", g_err.c_str());
        std::abort();
    }
}

void stepgrid_fluxes_dclaw_(double* q, double* aux, const int* mbc, const int* mx, const int* my,
                            const int* maux, const double* dx, const double* dy,
                            const double* dt, double* cflgrid, double* fm, double* fp,
                            double* gm, double* gp)
{
    double* const flux[4] = {fm, fp, gm, gp};
    int rc = (fm && fp && gm && gp) ? DCLAW_OK : fail(DCLAW_EINVAL, "null flux array");
    if (rc == DCLAW_OK)
        rc = stepgrid_host(q, aux, *mbc, *mx, *my, *maux, *dx, *dy, *dt, cflgrid, flux);
    if (rc != DCLAW_OK) {
        std::fprintf(stderr, "stepgrid_fluxes_dclaw: %s\n", g_err.c_str());
        std::abort();
    }
}

}  // extern "C"
