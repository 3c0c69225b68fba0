// swfd_kernels.hip -- CDNA4 (gfx950) kernels and the C ABI of libswfd.so.
//
// One fused kernel advances every shot of a batch by one leapfrog step:
// interior update, point / dense / bilinear-residual sources, Mur edges and
// corners, optional imaging condition, and (in spare workgroups) the bilinear
// receiver gather of the previous step's field. Fields are float32 (S, ny, nx)
// row-major, x fastest.
//
// Work decomposition: a workgroup owns a tile of 64 columns (one wavefront
// wide, so every row segment is a coalesced 256-byte access) by waves*RB rows
// and loops over a group of shots, keeping the medium coefficients of its
// cells in registers across the shot loop. Each lane walks a column strip of
// RB rows and reuses the vertical neighbours it already loaded. Tiles are
// numbered so that all shot groups of a tile, and neighbouring tiles, land on
// the same XCD (workgroup i runs on XCD i % 8), which keeps each XCD's L2
// serving the coefficients and the halo rows its tiles share.
//
// Boundary cells (outermost ring) are few; their lanes recompute the
// neighbouring interior values they depend on straight from the read-only
// inputs, so no workgroup ever waits for another and the output buffer never
// aliases an input.
//
// Arithmetic: the interior update uses fused multiply-adds in one fixed
// order shared by the main path and the boundary recomputation (so both see
// bit-identical interior values); the Mur update, the gather and the residual
// use separately rounded products and sums in the reference's op order.
// The library is compiled with -ffp-contract=off so that only the explicit
// __fmaf_rn calls fuse.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "swfd.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

#define SWFD_HIP_CHECK(expr)                                                            \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(SWFD_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kLanes = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kRows = 8;              // rows per lane strip
constexpr int kMaxImageGroups = 8;    // shot groups with separate imaging partials
constexpr int kXcds = 8;
constexpr int kGatherThreads = 256;

// All kernel inputs of one step. Passed by value (kernarg segment).
struct StepParams {
    int ny, nx, n_shots;
    float idx2, idy2, cen, dt2;
    const float *c2, *c2dt2, *mur;
    const float *f_old, *f_cur;
    float *f_out, *f_out2;
    // point source
    const int *src_rows, *src_cols;
    const float *src_values;
    int src_index;
    // bilinear source (NULL: point source): (S, 4) weights of the 2 x 2 cells
    // from (src_rows[s], src_cols[s]), order w00, w10, w01, w11
    const float *src_w;
    // dense source (S, ny, nx)
    const float *dense;
    // bilinear residual injection
    const float *resid;
    const int *cell_map, *entry_start, *pair_rec;
    const float *pair_w;
    int n_rec;
    // imaging
    const float *img, *g_in;
    float *g_out, *g_part;
    float img_scale;
    // gather of gat_field (S, ny, nx) into gat_out (S, R)
    const float *gat_field;
    float *gat_out;
    const int *rec_row0, *rec_col0;
    const float *rec_w;
    // launch geometry
    int tiles_x, n_groups, shots_per_group, waves;
    int n_work, chunk, n_stencil_blocks, n_edge_blocks;
    int rows, order;
    // two-step launches: outputs of the second step, a second gather, and
    // the marching geometry (column blocks of 62, row segments)
    float *f_out_b, *f_out2_b;
    const float *gat_field2;
    float *gat_out2;
    int n_cb, n_seg, seg;
    // two adjoint steps per launch: residual and imaging field of the second
    // step, shots per wavefront, rows of the imaging accumulator
    const float *resid_b, *img_b;
    int shots_per_wave, acc_rows;
    // full-width launches: wavefronts per workgroup, column offset of wave 0,
    // time levels per launch and each level's outputs (NULL: not stored)
    int wide_waves, wide_off;
    int levels;
    float *lv_out[4], *lv_out2[4];
    // Laplacian history: level j stores the Laplacian of its input field
    float *lv_lap[4];
    int hist_lap;
    // multi-step adjoint launches: each level's imaging field and residual
    const float *lv_in[4], *lv_resid[4];
    // receiver samples of each level of a multi-step forward launch (S, R)
    float *lv_samples[4];
    // per row: bit w set when full-width wavefront w holds a receiver
    // footprint cell there (NULL: intermediate levels stored everywhere)
    const unsigned *fp_rows;
    // boundary ring: the boundary cells of the forward field after step q in
    // slot q - bnd_first (forward launches write it, reconstructing adjoint
    // launches read it; NULL: none)
    float *bnd;
    int bnd_first, bnd_slots;
    // reconstructing adjoint launches (first step n = src_index): the forward
    // fields after steps n and n - 1, each level's reconstructed field (NULL:
    // not stored), the first step that images, and the rows of the staged
    // boundary columns
    const float *r_next, *r_cur;
    float *lv_rout[4];
    int img_from, bcol_rows;
    // reconstructing adjoint launches: segments away from rows 0 / H - 1 take
    // the lean marches (every level images and rebuilds a field of step >= 0,
    // segments of at least two rows, rows of at least two wavefronts)
    int adjr_lean;
    // full-width forward launches with footprint stores: segments away from
    // rows 1 / H - 1 take the lean marches (rows of at least two wavefronts)
    int fwdm_lean;
};

// Floats per shot in a slot of the boundary ring: row 0, row H - 1, then
// columns 0 and W - 1 of rows 1 .. H - 2.
__host__ __device__ __forceinline__ size_t bnd_floats(int H, int W)
{
    return 2 * (size_t)W + 2 * (size_t)(H - 2);
}

// 2u - u_old of the leapfrog update as one fused operation: 2u is exact, so
// the single rounding of fma(2, u, -u_old) equals that of the separate
// product and difference.
__device__ __forceinline__ float leap(float u, float u_old)
{
    return __fmaf_rn(2.0f, u, -u_old);
}

__device__ __forceinline__ float lap5(float w, float e, float s, float n, float c,
                                      float idx2, float idy2, float cen)
{
    return __fmaf_rn(cen, c, __fmaf_rn(idy2, s + n, idx2 * (w + e)));
}

// The transmitter's contribution to cell (rho, x) of shot s, given the
// source value `val` of the step: a point source adds val at (sr, sc); a
// bilinear source (p.src_w set) adds val * w at the 2 x 2 cells from (sr, sc)
// with the receivers' bilinear weights, so that source and receiver are
// adjoint (the reference's suggested fix for reciprocity, injecting
// w * pulse[n] on the four footprint nodes). The weights are read with scalar
// loads only on the footprint's two rows.
__device__ __forceinline__ float add_transmitter(const StepParams &p, int s, float v, int rho,
                                                 int x, int sr, int sc, float val)
{
    if (p.src_w == nullptr) return (rho == sr && x == sc) ? v + val : v;
    const int dy = rho - sr;
    if (dy != 0 && dy != 1) return v;
    typedef const float __attribute__((address_space(4))) kw;
    const kw *w = (const kw *)p.src_w + (size_t)s * 4 + 2 * dy;
    const float w0 = w[0], w1 = w[1];
    const int dx = x - sc;
    if (dx == 0) return v + val * w0;
    if (dx == 1) return v + val * w1;
    return v;
}

// Rows the transmitter of a step touches, beyond its first: 1 for a bilinear
// source, 0 for a point source.
__device__ __forceinline__ int transmitter_rows(const StepParams &p)
{
    return p.src_w != nullptr ? 1 : 0;
}

// Sources added to the interior value of cell (y, x) of shot s, in the order
// point (or bilinear), dense, residual scatter.
__device__ __forceinline__ float add_sources(const StepParams &p, int s, size_t soff,
                                             size_t ci, int y, int x, float v)
{
    if (p.src_rows != nullptr)
        v = add_transmitter(p, s, v, y, x, p.src_rows[s], p.src_cols[s], p.src_values[p.src_index]);
    if (p.dense != nullptr) v += p.dense[soff + ci];
    if (p.resid != nullptr) {
        const int e = p.cell_map[ci];
        if (e >= 0) {
            const float *res = p.resid + (size_t)s * p.n_rec;
            float acc = 0.0f;
            for (int k = p.entry_start[e]; k < p.entry_start[e + 1]; ++k)
                acc += res[p.pair_rec[k]] * p.pair_w[k];
            v += acc;
        }
    }
    return v;
}

// Interior update of one cell read straight from global memory (used for the
// neighbours of boundary cells). Same arithmetic as the main path.
template <bool ADJ>
__device__ float interior_value(const StepParams &p, int s, size_t soff, int y, int x)
{
    const int W = p.nx, H = p.ny;
    const float *u = p.f_cur + soff;
    const size_t i = (size_t)y * W + x;
    const float uc = u[i];
    const float uw = x > 0 ? u[i - 1] : 0.0f;
    const float ue = x < W - 1 ? u[i + 1] : 0.0f;
    const float us = y > 0 ? u[i - W] : 0.0f;
    const float un = y < H - 1 ? u[i + W] : 0.0f;
    const float uo = p.f_old[soff + i];
    float v;
    if (!ADJ) {
        v = __fmaf_rn(p.c2dt2[i], lap5(uw, ue, us, un, uc, p.idx2, p.idy2, p.cen),
                      leap(uc, uo));
    } else {
        const float *c2 = p.c2;
        const float pw = x > 0 ? c2[i - 1] * uw : 0.0f;
        const float pe = x < W - 1 ? c2[i + 1] * ue : 0.0f;
        const float ps = y > 0 ? c2[i - W] * us : 0.0f;
        const float pn = y < H - 1 ? c2[i + W] * un : 0.0f;
        v = __fmaf_rn(p.dt2, lap5(pw, pe, ps, pn, c2[i] * uc, p.idx2, p.idy2, p.cen),
                      leap(uc, uo));
    }
    return add_sources(p, s, soff, i, y, x, v);
}

// Mur update of an edge cell that is not a corner (solver.py:99-109).
template <bool ADJ>
__device__ float edge_value(const StepParams &p, int s, size_t soff, int y, int x)
{
    const int W = p.nx, H = p.ny;
    const float *u = p.f_cur + soff;
    if (x == 0) {
        const float k = p.mur[y];
        return u[(size_t)y * W + 1] +
               k * (interior_value<ADJ>(p, s, soff, y, 1) - u[(size_t)y * W]);
    }
    if (x == W - 1) {
        const float k = p.mur[H + y];
        return u[(size_t)y * W + W - 2] +
               k * (interior_value<ADJ>(p, s, soff, y, W - 2) - u[(size_t)y * W + W - 1]);
    }
    if (y == 0) {
        const float k = p.mur[2 * H + x];
        return u[(size_t)W + x] + k * (interior_value<ADJ>(p, s, soff, 1, x) - u[x]);
    }
    const float k = p.mur[2 * H + W + x];
    return u[(size_t)(H - 2) * W + x] +
           k * (interior_value<ADJ>(p, s, soff, H - 2, x) - u[(size_t)(H - 1) * W + x]);
}

// Mur update of any boundary cell; corners average an x- and a y-update that
// read the updated edge cells next to them (solver.py:111-124).
template <bool ADJ>
__device__ float boundary_value(const StepParams &p, int s, size_t soff, int y, int x)
{
    const int W = p.nx, H = p.ny;
    const bool ex = (x == 0) || (x == W - 1);
    const bool ey = (y == 0) || (y == H - 1);
    if (!(ex && ey)) return edge_value<ADJ>(p, s, soff, y, x);
    const float *u = p.f_cur + soff;
    const int ci = (y == 0 ? 0 : 2) + (x == 0 ? 0 : 1);
    const int r_in = (y == 0) ? 1 : H - 2;
    const int c_in = (x == 0) ? 1 : W - 2;
    const float uc = u[(size_t)y * W + x];
    const float *kx = p.mur + 2 * H + 2 * W;
    const float *ky = kx + 4;
    const float from_x =
        u[(size_t)y * W + c_in] + kx[ci] * (edge_value<ADJ>(p, s, soff, y, c_in) - uc);
    const float from_y =
        u[(size_t)r_in * W + x] + ky[ci] * (edge_value<ADJ>(p, s, soff, r_in, x) - uc);
    return 0.5f * (from_x + from_y);
}

// Bilinear gather of (s, r) pairs, in the reference's product/sum order.
__device__ __forceinline__ void gather_pairs(const StepParams &p, int t)
{
    const int R = p.n_rec;
    const int total = p.n_shots * R;
    const float *field = p.gat_field;
    float *dest = p.gat_out;
    if (t >= total) {
        t -= total;
        if (p.gat_field2 == nullptr || t >= total) return;
        field = p.gat_field2;
        dest = p.gat_out2;
    }
    const int s = t / R;
    const int r = t - s * R;
    const int W = p.nx;
    const float *f = field + (size_t)s * ((size_t)p.ny * W);
    const size_t i = (size_t)p.rec_row0[r] * W + p.rec_col0[r];
    const float *w = p.rec_w;
    float acc = f[i] * w[r];
    acc = acc + f[i + 1] * w[R + r];
    acc = acc + f[i + W] * w[2 * R + r];
    acc = acc + f[i + W + 1] * w[3 * R + r];
    dest[t] = acc;
}

// Boundary ring of one (shot, cell) item: item t covers shot t / NB and the
// (t % NB)-th boundary cell, NB = 2 (nx + ny) - 4, ordered bottom row, top
// row, left column, right column. Kept out of line and called before any
// stencil state is live, so it does not raise the stencil's register budget.
template <bool ADJ>
__device__ __forceinline__ void edge_work(const StepParams &p, int t)
{
    const int W = p.nx, H = p.ny;
    const int nb = 2 * (W + H) - 4;
    if (t >= nb * p.n_shots) return;
    const int s = t / nb;
    int i = t - s * nb;
    int y, x;
    if (i < W) {
        y = 0;
        x = i;
    } else if (i < 2 * W) {
        y = H - 1;
        x = i - W;
    } else if (i < 2 * W + H - 2) {
        x = 0;
        y = 1 + (i - 2 * W);
    } else {
        x = W - 1;
        y = 1 + (i - 2 * W - (H - 2));
    }
    const size_t soff = (size_t)s * H * W;
    const float v = boundary_value<ADJ>(p, s, soff, y, x);
    const size_t ci = soff + (size_t)y * W + x;
    p.f_out[ci] = v;
    if (p.f_out2 != nullptr) p.f_out2[ci] = v;
}

// Feature flags of a specialised step kernel; the host derives them from the
// non-null pointers of StepParams, so the generic boundary path (which tests
// the pointers at run time) always agrees with the specialised main path.
enum : unsigned {
    kPoint = 1u,
    kDense = 2u,
    kResid = 4u,
    kImage = 8u,
    kOut2 = 16u,
    kGather = 32u,
    kLap = 64u,  // multi-step forward: store each level's input Laplacian
    kBnd = 128u, // multi-step forward: footprint stores keep the boundary cells too
};

// Neighbour exchange inside a wavefront through DPP wavefront shifts: lane i
// receives lane i-1 (west) or lane i+1 (east); the lane at the wavefront's
// edge receives `halo` instead. Every lane must execute these (no divergence).
__device__ __forceinline__ float from_west(float v, float halo)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(halo), __float_as_int(v),
                                                      0x138 /* wave_shr:1 */, 0xf, 0xf, false));
}

__device__ __forceinline__ float from_east(float v, float halo)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(halo), __float_as_int(v),
                                                      0x130 /* wave_shl:1 */, 0xf, 0xf, false));
}

// Pointers in the global address space: loads become global_load with a
// scalar (per-wavefront) row base and one 32-bit lane offset, instead of flat
// loads with a 64-bit address per lane and row.
#define SWFD_GLOBAL __attribute__((address_space(1)))
typedef const float SWFD_GLOBAL gcf;
typedef float SWFD_GLOBAL gf;
typedef const int SWFD_GLOBAL gci;

__device__ __forceinline__ gcf *gload_ptr(const float *p) { return (gcf *)p; }

// Raw buffer view of `bytes` bytes at p (word 3 = 0x00020000 for gfx9-family
// raw 32-bit access), and a 32-bit load at lane byte offset voff plus
// wavefront-uniform byte offset soff.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const float *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ gf *gstore_ptr(float *p) { return (gf *)p; }

// 32-bit load at a lane byte offset from a wavefront-uniform base (a global
// load with the base in SGPRs and a 32-bit lane offset).
__device__ __forceinline__ float gload_off(const float *base, int voff)
{
    return *(const gcf *)((const char SWFD_GLOBAL *)base + (unsigned)voff);
}

// Work-item order (p.order): 0 = strip, column group, shot group (fastest);
// 1 = shot group, strip, column group (fastest): a workgroup's wavefronts
// cover adjacent columns and consecutive strips follow each other;
// 2 = strip, shot group, column group (fastest).
__device__ __forceinline__ void decode_item(const StepParams &p, int item, int &grp, int &cg,
                                            int &strip)
{
    const int strips = (p.ny + p.rows - 1) / p.rows;
    if (p.order == 1) {
        cg = item % p.tiles_x;
        const int t = item / p.tiles_x;
        strip = t % strips;
        grp = t / strips;
    } else if (p.order == 2) {
        cg = item % p.tiles_x;
        const int t = item / p.tiles_x;
        grp = t % p.n_groups;
        strip = t / p.n_groups;
    } else {
        grp = item % p.n_groups;
        const int t = item / p.n_groups;
        cg = t % p.tiles_x;
        strip = t / p.tiles_x;
    }
}

// One wavefront = one work item: 64 columns x RB rows x a group of shots.
// Lane l owns column x0 + l and walks the RB rows of the strip; vertical
// neighbours come from its own registers, horizontal ones from the adjacent
// lanes by DPP, and only lanes 0 and 63 load the column just outside the
// strip. Coefficients of the strip stay in registers across the shot loop.
// All indices except the lane's column are uniform over the wavefront, so row
// addresses live in scalar registers. Boundary cells are skipped here and
// written by the edge workgroups. Imaging (adjoint) adds the shot group's
// partial sum of the previous step's lam * Lap(forward field) into its own
// slice g_part[group].
template <bool ADJ, int RB, unsigned F>
__global__ void __launch_bounds__(256) step_kernel(StepParams p)
{
    if ((int)blockIdx.x < p.n_edge_blocks) {
        edge_work<ADJ>(p, (int)(blockIdx.x * blockDim.x + threadIdx.x));
        return;
    }
    const int b = blockIdx.x - p.n_edge_blocks;
    if constexpr ((F & kGather) != 0) {
        if (b >= p.n_stencil_blocks) {
            gather_pairs(p, (b - p.n_stencil_blocks) * (int)blockDim.x + (int)threadIdx.x);
            return;
        }
    }
    // XCD-aware numbering: consecutive work items stay on one XCD.
    const int wblk = (b % kXcds) * p.chunk + b / kXcds;
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kLanes));
    const int item = wblk * kWavesPerBlock + wave;
    if (item >= p.n_work) return;  // uniform over the wavefront
    int grp, cg, strip;
    decode_item(p, item, grp, cg, strip);
    const int W = p.nx, H = p.ny;
    const int x0 = cg * kLanes;
    const int x = x0 + lane;
    const int y0 = strip * RB;
    const bool xin = x < W;
    const int hx = lane == 0 ? x0 - 1 : x0 + kLanes;
    const bool hneed = ((lane == 0) & (x0 > 0)) | ((lane == kLanes - 1) & (x0 + kLanes < W));
    const size_t plane = (size_t)H * W;

    // Loads are buffer loads at clamped offsets (always inside the field):
    // one lane-offset VGPR for every row, row offsets in scalar registers,
    // and out-of-domain neighbours zeroed by a select afterwards.
    const unsigned plane_bytes = (unsigned)(plane * sizeof(float));
    const int xo = min(x, W - 1) * 4;
    const int ho = min(max(hx, 0), W - 1) * 4;
    auto row_off = [&](int yy) -> int { return min(max(yy, 0), H - 1) * W * 4; };

    float coef[RB];  // dt^2 c^2 (forward) or c^2 (adjoint) of the lane's cells
    float c2v[2];    // adjoint: c^2 on the rows just above and below the strip
    float c2h[RB];   // adjoint: c^2 on the halo column of lanes 0 / 63
    {
        const __amdgpu_buffer_rsrc_t cf = make_rsrc(ADJ ? p.c2 : p.c2dt2, plane_bytes);
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const int yy = y0 + i;
            const float a = bload(cf, xo, row_off(yy));
            coef[i] = (xin & (yy < H)) ? a : 0.0f;
            if constexpr (ADJ) {
                const float h = bload(cf, ho, row_off(yy));
                c2h[i] = (hneed & (yy < H)) ? h : 0.0f;
            }
        }
        if constexpr (ADJ) {
            const float a = bload(cf, xo, row_off(y0 - 1));
            const float b = bload(cf, xo, row_off(y0 + RB));
            c2v[0] = (xin & (y0 >= 1)) ? a : 0.0f;
            c2v[1] = (xin & (y0 + RB < H)) ? b : 0.0f;
        }
    }
    bool inject = false;
    if constexpr ((F & kResid) != 0) {
        const __amdgpu_buffer_rsrc_t cm = make_rsrc((const float *)p.cell_map, plane_bytes);
        bool mine = false;
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const int yy = y0 + i;
            const int e = __float_as_int(bload(cm, xo, row_off(yy)));
            mine |= xin & (yy < H) & (e >= 0);
        }
        inject = __any(mine);
    }
    float gacc[RB];
#pragma unroll
    for (int i = 0; i < RB; ++i) gacc[i] = 0.0f;

    const int s_begin = grp * p.shots_per_group;
    const int s_end = min(p.n_shots, s_begin + p.shots_per_group);
    for (int s = s_begin; s < s_end; ++s) {
        const size_t soff = (size_t)s * plane;
        const __amdgpu_buffer_rsrc_t ru = make_rsrc(p.f_cur + soff, plane_bytes);
        const __amdgpu_buffer_rsrc_t ro = make_rsrc(p.f_old + soff, plane_bytes);
        float col[RB + 2], hal[RB], old[RB];
#pragma unroll
        for (int i = 0; i < RB + 2; ++i) {
            const int yy = y0 - 1 + i;
            const float a = bload(ru, xo, row_off(yy));
            col[i] = (xin & (yy >= 0) & (yy < H)) ? a : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const int yy = y0 + i;
            const float h = bload(ru, ho, row_off(yy));
            const float o = bload(ro, xo, row_off(yy));
            hal[i] = (hneed & (yy < H)) ? h : 0.0f;
            old[i] = (xin & (yy < H)) ? o : 0.0f;
        }
        // forward field of the imaging condition, loaded with the rest
        float fc[RB + 2], fh[RB];
        if constexpr (ADJ && (F & kImage) != 0) {
            const __amdgpu_buffer_rsrc_t rf = make_rsrc(p.img + soff, plane_bytes);
#pragma unroll
            for (int i = 0; i < RB + 2; ++i) {
                const int yy = y0 - 1 + i;
                const float a = bload(rf, xo, row_off(yy));
                fc[i] = (xin & (yy >= 0) & (yy < H)) ? a : 0.0f;
            }
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                const int yy = y0 + i;
                const float h = bload(rf, ho, row_off(yy));
                fh[i] = (hneed & (yy < H)) ? h : 0.0f;
            }
        }
        float v[RB];
        if constexpr (!ADJ) {
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                const float uw = from_west(col[i + 1], hal[i]);
                const float ue = from_east(col[i + 1], hal[i]);
                v[i] = __fmaf_rn(coef[i],
                                 lap5(uw, ue, col[i], col[i + 2], col[i + 1], p.idx2, p.idy2,
                                      p.cen),
                                 leap(col[i + 1], old[i]));
            }
        } else {
            float pc[RB + 2];
            pc[0] = c2v[0] * col[0];
            pc[RB + 1] = c2v[1] * col[RB + 1];
#pragma unroll
            for (int i = 0; i < RB; ++i) pc[i + 1] = coef[i] * col[i + 1];
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                const float ph = c2h[i] * hal[i];
                const float pw = from_west(pc[i + 1], ph);
                const float pe = from_east(pc[i + 1], ph);
                v[i] = __fmaf_rn(p.dt2,
                                 lap5(pw, pe, pc[i], pc[i + 2], pc[i + 1], p.idx2, p.idy2, p.cen),
                                 leap(col[i + 1], old[i]));
            }
        }
        if constexpr ((F & kPoint) != 0) {
            const int sr = p.src_rows[s], sc = p.src_cols[s];
            const int extra = transmitter_rows(p);
            if (sr + extra >= y0 && sr < y0 + RB && (unsigned)(x - sc) <= (unsigned)extra) {
                const float val = p.src_values[p.src_index];
#pragma unroll
                for (int i = 0; i < RB; ++i) v[i] = add_transmitter(p, s, v[i], y0 + i, x, sr, sc, val);
            }
        }
        if constexpr ((F & kDense) != 0) {
            const __amdgpu_buffer_rsrc_t rd = make_rsrc(p.dense + soff, plane_bytes);
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                const int yy = y0 + i;
                const float d = bload(rd, xo, row_off(yy));
                v[i] += (xin & (yy < H)) ? d : 0.0f;
            }
        }
        if constexpr ((F & kResid) != 0) {
            if (inject) {
                const float *res = p.resid + (size_t)s * p.n_rec;
#pragma unroll
                for (int i = 0; i < RB; ++i) {
                    const int yy = y0 + i;
                    const int e = (xin & (yy < H)) ? p.cell_map[(size_t)yy * W + x] : -1;
                    if (e >= 0) {
                        float acc = 0.0f;
                        for (int k = p.entry_start[e]; k < p.entry_start[e + 1]; ++k)
                            acc += res[p.pair_rec[k]] * p.pair_w[k];
                        v[i] += acc;
                    }
                }
            }
        }
        gf *out = gstore_ptr(p.f_out) + soff;
        const bool xint = xin && x > 0 && x < W - 1;
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const int yy = y0 + i;
            if (xint && yy > 0 && yy < H - 1) {
                const size_t ci = (size_t)yy * W + x;
                out[ci] = v[i];
                if constexpr ((F & kOut2) != 0) (gstore_ptr(p.f_out2) + soff)[ci] = v[i];
            }
        }
        // Imaging of the previous (later-in-time) adjoint step: its field is
        // this step's input u (already Mur-updated), held in col[].
        if constexpr (ADJ && (F & kImage) != 0) {
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                const float fw = from_west(fc[i + 1], fh[i]);
                const float fe = from_east(fc[i + 1], fh[i]);
                const float lf = lap5(fw, fe, fc[i], fc[i + 2], fc[i + 1], p.idx2, p.idy2, p.cen);
                gacc[i] = __fmaf_rn(col[i + 1], lf, gacc[i]);
            }
        }
    }
    if constexpr (ADJ && (F & kImage) != 0) {
        gf *gp = gstore_ptr(p.g_part) + (size_t)grp * plane;
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const int yy = y0 + i;
            if (xin & (yy < H)) {
                gf *cell = gp + (size_t)yy * W + x;
                *cell = __fmaf_rn(p.img_scale, gacc[i], *cell);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Two forward steps per launch (temporal blocking)
// ---------------------------------------------------------------------------
//
// A wavefront covers 64 columns [xs, xs+64), one per lane, and a segment of
// rows [r0, r1), and marches down the rows for each shot of its group. The
// first step (u^{n+1} = C from u^{n-1} = A, u^n = B) runs one row ahead of the
// second (u^{n+2} = D from B, C), so each B, A and coefficient row is loaded
// once, and per two steps the kernel reads two fields and writes two, in whole
// 256-byte rows. The second step needs C one column beyond each side of the
// block: lanes 0 and 63 compute those two halo columns as well (same
// arithmetic, neighbours from their own registers and one extra load), with
// their outer B neighbours loaded by lanes 1 and 62. Mur edges, corners and
// point sources are applied in the march with exactly the single-step
// kernel's arithmetic, so results are bit-identical to two single steps; the
// top row, whose Mur update needs the row below, is finalised one row later.
// Blocks start at multiples of 64, except that the last block ends at column
// W-1 (it may overlap the previous block and stores only its own columns), so
// column W-1 is never on lane 0 and its Mur update finds its west neighbour.

constexpr int kSegRows = 64;         // rows a two-step wavefront marches (at least)

// Rows [r0, r1) of segment `seg`: n_seg near-equal parts of H, each of at
// least p.seg >= 2 rows (the top row's Mur update in the second step reads
// the row below it, so a segment must hold both).
__device__ __forceinline__ void seg_rows(const StepParams &p, int seg, int &r0, int &r1)
{
    r0 = (int)(((long)seg * p.ny) / p.n_seg);
    r1 = (int)(((long)(seg + 1) * p.ny) / p.n_seg);
}
#ifndef SWFD_PREFETCH_ROWS
#define SWFD_PREFETCH_ROWS 2
#endif
constexpr int kPf = SWFD_PREFETCH_ROWS;  // rows loaded ahead of the march

// Problem of workgroup b in a batched launch: entry e owns workgroups
// [first[e], first[e + 1]) (a wavefront-uniform scan of a short table).
__device__ __forceinline__ int batch_entry(const int *first, int n, int b)
{
    int e = 0;
    while (e + 1 < n && first[e + 1] <= b) ++e;
    return e;
}

// Residual injection for the lanes of a wavefront whose cell has scatter
// entry e >= 0 (others pass e = -1): lane by lane over the flagged lanes, with
// the entry's (receiver, weight) pairs and the residuals read through scalar
// loads (wavefront-uniform addresses), so no wait on the vector memory
// pipeline (which holds the march's prefetched rows) is needed. Per cell the
// sum is the single-step kernel's: a = 0, a += res[rec_k] * w_k in entry order.
typedef const int __attribute__((address_space(4))) kci;    // read-only, scalar-loadable
typedef const float __attribute__((address_space(4))) kcf;
typedef const StepParams __attribute__((address_space(4))) kparams;

// Entry e of a batched launch's parameter table, read through the constant
// address space: the table does not change while the kernel runs, so its
// (wavefront-uniform) fields load into scalar registers like kernel arguments
// instead of occupying vector registers.
__device__ __forceinline__ const StepParams &batch_params(const StepParams *table, int e)
{
    const kparams *t = (const kparams *)table;
    return *(const StepParams *)(t + e);
}

__device__ __forceinline__ float inject_lanes(const StepParams &p, const float *res,
                                              size_t res_off, int e, int lane, float v)
{
    unsigned long long m = __ballot(e >= 0);
    if (m == 0) return v;
    const kci *start = (const kci *)p.entry_start;
    const kci *rec = (const kci *)p.pair_rec;
    const kcf *w = (const kcf *)p.pair_w;
    const kcf *r = (const kcf *)res + res_off;
    do {
        const int l = (int)__builtin_ctzll(m);
        m &= m - 1;
        const int eu = __builtin_amdgcn_readlane(e, l);
        float a = 0.0f;
        const int k1 = start[eu + 1];
        for (int k = start[eu]; k < k1; ++k) a += r[rec[k]] * w[k];
        v = lane == l ? v + a : v;
    } while (m != 0);
    return v;
}

// Mur update of a top/bottom edge cell (x in [1, W-2]) and of the corners of
// that row, from the row's current-time values `cur_r`, the interior-side
// row's current and next values `cur_in`, `nxt_in`, and the side edge values
// of the interior-side row (`nxt_in` at x = 0 / W-1 already Mur-updated).
// `ci` selects corners 0/1 (top) or 2/3 (bottom). Every lane must call it.
__device__ __forceinline__ float mur_row(const float *mur, int W, int H, int x, float k_row_x,
                                         float cur_r, float cur_in, float nxt_in, int ci)
{
    // edge cells: u'[r][x] = u[in][x] + k (u'[in][x] - u[r][x])
    float v = cur_in + k_row_x * (nxt_in - cur_r);
    const float ve = from_east(v, 0.0f), vw = from_west(v, 0.0f);
    const float ce = from_east(cur_r, 0.0f), cw = from_west(cur_r, 0.0f);
    const float *kx = mur + 2 * H + 2 * W;
    const float *ky = kx + 4;
    if (x == 0) {
        const float from_x = ce + kx[ci] * (ve - cur_r);
        const float from_y = cur_in + ky[ci] * (nxt_in - cur_r);
        v = 0.5f * (from_x + from_y);
    } else if (x == W - 1) {
        const float from_x = cw + kx[ci + 1] * (vw - cur_r);
        const float from_y = cur_in + ky[ci + 1] * (nxt_in - cur_r);
        v = 0.5f * (from_x + from_y);
    }
    return v;
}

// Mur update of the left / right edge cells of an interior row (1 <= y <= H-2);
// k_side is the row's left coefficient on column 0 and its right coefficient
// on column W-1 (other lanes ignore it).
__device__ __forceinline__ float mur_sides(float v, float cur, int x, int W, float k_side)
{
    const float ve = from_east(v, 0.0f), vw = from_west(v, 0.0f);
    const float ce = from_east(cur, 0.0f), cw = from_west(cur, 0.0f);
    if (x == 0) return ce + k_side * (ve - cur);
    if (x == W - 1) return cw + k_side * (vw - cur);
    return v;
}

template <bool EDGE, unsigned F>
__device__ __forceinline__ void fwd2_wave(const StepParams &p, int lane, int grp, int cb, int seg)
{
    const int W = p.nx, H = p.ny;
    const int own0 = cb * kLanes;
    const int xs = min(own0, max(W - kLanes, 0));
    const int x = xs + lane;
    const bool xin = x < W;
    const bool owned = xin & (x >= own0);
    const int xo = min(x, W - 1) * 4;
    // halo columns (lane 0: xs-1, lane 63: xs+64) and the B column loaded
    // beside them (lane 1: xs-2, lane 62: xs+65); other lanes use their own
    const bool hl = lane == 0, hr = lane == kLanes - 1;
    const int xh = hl ? xs - 1 : (hr ? xs + kLanes : x);
    const int xq = lane == 1 ? xs - 2 : (lane == kLanes - 2 ? xs + kLanes + 1 : xh);
    const bool xh_in = (xh >= 0) & (xh < W);
    const bool xq_in = (xq >= 0) & (xq < W);
    const int ho = min(max(xh, 0), W - 1) * 4;
    const int qo = min(max(xq, 0), W - 1) * 4;
    int r0, r1;
    seg_rows(p, seg, r0, r1);
    const int y_start = r0 == 0 ? 0 : r0 - 1;
    const size_t plane = (size_t)H * W;
    const unsigned plane_bytes = (unsigned)(plane * sizeof(float));
    auto row_off = [&](int yy) -> int { return min(max(yy, 0), H - 1) * W * 4; };
    const __amdgpu_buffer_rsrc_t rk = make_rsrc(p.c2dt2, plane_bytes);
    const float *mur = p.mur;
    // per-row side coefficients: mur[y] (left) for column 0, mur[H + y]
    // (right) for the other lanes, read as a row slot like the fields
    const __amdgpu_buffer_rsrc_t rm = make_rsrc(p.mur, (unsigned)(2 * H * sizeof(float)));
    const int mo = (x == 0) ? 0 : H * 4;
    auto mrow = [&](int yy) -> int { return min(max(yy, 0), H - 1) * 4; };
    float kb_x = 0.0f, kt_x = 0.0f, kb_h = 0.0f, kt_h = 0.0f;
    // corner coefficients of column W-1 for the right halo
    float kx1 = 0.0f, ky1 = 0.0f, kx3 = 0.0f, ky3 = 0.0f;
    const bool h_right_edge = hr & (xh == W - 1);
    if constexpr (EDGE) {
        kb_x = mur[2 * H + xo / 4];
        kt_x = mur[2 * H + W + xo / 4];
        kb_h = mur[2 * H + ho / 4];
        kt_h = mur[2 * H + W + ho / 4];
        const float *kx = mur + 2 * H + 2 * W;
        kx1 = kx[1];
        ky1 = kx[4 + 1];
        kx3 = kx[3];
        ky3 = kx[4 + 3];
    }
    const int s_begin = grp * p.shots_per_group;
    const int s_end = min(p.n_shots, s_begin + p.shots_per_group);
    for (int s = s_begin; s < s_end; ++s) {
        const size_t soff = (size_t)s * plane;
        const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.f_old + soff, plane_bytes);
        const __amdgpu_buffer_rsrc_t rb = make_rsrc(p.f_cur + soff, plane_bytes);
        gf *outc = gstore_ptr(p.f_out) + soff;
        gf *outd = gstore_ptr(p.f_out_b) + soff;
        int sr = -1, sc = -1;
        float s1 = 0.0f, s2 = 0.0f;
        if constexpr ((F & kPoint) != 0) {
            sr = p.src_rows[s];
            sc = p.src_cols[s];
            s1 = p.src_values[p.src_index];
            s2 = p.src_values[p.src_index + 1];
        }
        float bm, b0, qm1, q0;
        {
            const float t0 = bload(rb, xo, row_off(y_start - 1));
            const float t1 = bload(rb, xo, row_off(y_start));
            const float t2 = bload(rb, qo, row_off(y_start - 1));
            const float t3 = bload(rb, qo, row_off(y_start));
            bm = (xin & (y_start >= 1)) ? t0 : 0.0f;
            b0 = xin ? t1 : 0.0f;
            qm1 = (xq_in & (y_start >= 1)) ? t2 : 0.0f;
            q0 = xq_in ? t3 : 0.0f;
        }
        // Rows arrive kPf iterations ahead in kPf register slots used in
        // turn: a slot is consumed and refilled in place, so no load in
        // flight is copied and the wait for it is never forced early. Slot
        // loads: B at row y+1 (block and halo columns), A and the
        // coefficients at row y (block and halo columns). Row offsets are
        // clamped, so loads past the last row read valid memory; their
        // values are masked. (Slots are issued in order, as in the steady
        // state, so the first row waits only for its own slot.)
        float qb[kPf], qq[kPf], qa[kPf], qk[kPf], qha[kPf], qhk[kPf], qm[kPf];
#pragma unroll
        for (int i = 0; i < kPf; ++i) {
            __builtin_amdgcn_sched_barrier(0);
            qb[i] = bload(rb, xo, row_off(y_start + 1 + i));
            qq[i] = bload(rb, qo, row_off(y_start + 1 + i));
            qa[i] = bload(ra, xo, row_off(y_start + i));
            qk[i] = bload(rk, xo, row_off(y_start + i));
            qha[i] = bload(ra, ho, row_off(y_start + i));
            qhk[i] = bload(rk, ho, row_off(y_start + i));
            qm[i] = 0.0f;
            if constexpr (EDGE) qm[i] = bload(rm, mo, mrow(y_start + i));
        }
        __builtin_amdgcn_sched_barrier(0);
        float cm2 = 0.0f, cm1 = 0.0f, km1 = 0.0f, dm1 = 0.0f, sm1 = 0.0f, hcm1 = 0.0f;
        auto row = [&](int y, float &sb, float &sq, float &sa, float &sk, float &sha, float &shk,
                       float &sm) {
            const bool row_in = y < H;
            float bp = (xin & (y + 1 < H)) ? sb : 0.0f;
            float qp = (xq_in & (y + 1 < H)) ? sq : 0.0f;
            float a0 = (xin & row_in) ? sa : 0.0f;
            float k0 = (xin & row_in) ? sk : 0.0f;
            float ha0 = (xh_in & row_in) ? sha : 0.0f;
            float hk0 = (xh_in & row_in) ? shk : 0.0f;
            float s0 = sm;
            // Materialise the values here (empty asm, no instruction) and
            // keep the refill after them, so each slot's register is reused
            // instead of copied on the back edge.
            asm volatile("" : "+v"(bp), "+v"(qp), "+v"(a0), "+v"(k0), "+v"(ha0), "+v"(hk0),
                         "+v"(s0));
            __builtin_amdgcn_sched_barrier(0);
            // rows past r1 are not needed: reload row r1 instead (a cache hit;
            // a conditional refill would make every wait a full one)
            const int yf = min(y + kPf, r1);
            sb = bload(rb, xo, row_off(yf + 1));
            sq = bload(rb, qo, row_off(yf + 1));
            sa = bload(ra, xo, row_off(yf));
            sk = bload(rk, xo, row_off(yf));
            sha = bload(ra, ho, row_off(yf));
            shk = bload(rk, ho, row_off(yf));
            if constexpr (EDGE) sm = bload(rm, mo, mrow(yf));
            // ---- first step, row y: block columns, then the halo columns
            float c0 = 0.0f, hc = 0.0f;
            if (row_in) {
                const float bw = from_west(b0, q0), be = from_east(b0, q0);
                c0 = __fmaf_rn(k0, lap5(bw, be, bm, bp, b0, p.idx2, p.idy2, p.cen),
                               leap(b0, a0));
                if constexpr ((F & kPoint) != 0) c0 = add_transmitter(p, s, c0, y, x, sr, sc, s1);
                if constexpr (EDGE) {
                    if (y >= 1 && y <= H - 2) c0 = mur_sides(c0, b0, x, W, s0);
                    if (y == H - 1) c0 = mur_row(mur, W, H, x, kt_x, b0, bm, cm1, 2);
                }
                // halo: lane 0 has B(xs) as its east and B(xs-2) (lane 1) as
                // its west neighbour; lane 63 has B(xs+63) as its west and
                // B(xs+65) (lane 62) as its east neighbour
                const float qe = from_east(q0, 0.0f), qw = from_west(q0, 0.0f);
                const float hw = hl ? qe : b0;
                const float he = hl ? b0 : qw;
                hc = __fmaf_rn(hk0, lap5(hw, he, qm1, qp, q0, p.idx2, p.idy2, p.cen),
                               leap(q0, ha0));
                if constexpr ((F & kPoint) != 0) hc = add_transmitter(p, s, hc, y, xh, sr, sc, s1);
                if constexpr (EDGE) {
                    // right halo on column W-1: Mur from lane 63's own block
                    // values at column W-2 (b0, c0); never the left halo
                    if (h_right_edge) {
                        if (y >= 1 && y <= H - 2) {
                            hc = b0 + s0 * (c0 - q0);
                        } else if (y == H - 1) {
                            const float from_x = b0 + kx3 * (c0 - q0);
                            const float from_y = qm1 + ky3 * (hcm1 - q0);
                            hc = 0.5f * (from_x + from_y);
                        }
                    } else if (y == H - 1) {
                        hc = qm1 + kt_h * (hcm1 - q0);
                    }
                    if (y == 1) {
                        // row 0 (bottom edge) now that row 1 is final
                        cm1 = mur_row(mur, W, H, x, kb_x, bm, b0, c0, 0);
                        if (h_right_edge) {
                            const float from_x = bm + kx1 * (cm1 - qm1);
                            const float from_y = q0 + ky1 * (hc - qm1);
                            hcm1 = 0.5f * (from_x + from_y);
                        } else {
                            hcm1 = q0 + kb_h * (hc - qm1);
                        }
                        if (owned & (r0 == 0)) {
                            outc[xo / 4] = cm1;
                            if constexpr ((F & kOut2) != 0)
                                if (p.f_out2 != nullptr) (gstore_ptr(p.f_out2) + soff)[xo / 4] = cm1;
                        }
                    }
                }
                if (owned & (y >= r0) & (y < r1) & (y > 0)) {
                    outc[(size_t)y * W + x] = c0;
                    if constexpr ((F & kOut2) != 0)
                        if (p.f_out2 != nullptr) (gstore_ptr(p.f_out2) + soff)[(size_t)y * W + x] = c0;
                }
            }
            // ---- second step, row y - 1
            const int r = y - 1;
            float d = 0.0f;
            if (r >= r0 && r < r1) {
                const float cw = from_west(cm1, hcm1), ce = from_east(cm1, hcm1);
                d = __fmaf_rn(km1, lap5(cw, ce, cm2, c0, cm1, p.idx2, p.idy2, p.cen),
                              leap(cm1, bm));
                if constexpr ((F & kPoint) != 0) d = add_transmitter(p, s, d, r, x, sr, sc, s2);
                if constexpr (EDGE) {
                    if (r >= 1 && r <= H - 2) d = mur_sides(d, cm1, x, W, sm1);
                    if (r == H - 1) d = mur_row(mur, W, H, x, kt_x, cm1, cm2, dm1, 2);
                    if (r == 1) {
                        const float d0 = mur_row(mur, W, H, x, kb_x, cm2, cm1, d, 0);
                        if (owned) {
                            outd[xo / 4] = d0;
                            if constexpr ((F & kOut2) != 0)
                                if (p.f_out2_b != nullptr) (gstore_ptr(p.f_out2_b) + soff)[xo / 4] = d0;
                        }
                    }
                }
                if (owned & (r > 0)) {
                    outd[(size_t)r * W + x] = d;
                    if constexpr ((F & kOut2) != 0)
                        if (p.f_out2_b != nullptr) (gstore_ptr(p.f_out2_b) + soff)[(size_t)r * W + x] = d;
                }
            }
            cm2 = cm1;
            cm1 = c0;
            hcm1 = hc;
            bm = b0;
            b0 = bp;
            qm1 = q0;
            q0 = qp;
            km1 = k0;
            dm1 = d;
            sm1 = s0;
        };
        // rows in groups of kPf; rows past r1 store nothing (both steps'
        // stores are bounded by r1) and read clamped rows
        for (int y = y_start; y <= r1; y += kPf) {
#pragma unroll
            for (int i = 0; i < kPf; ++i)
                row(y + i, qb[i], qq[i], qa[i], qk[i], qha[i], qhk[i], qm[i]);
        }
    }
}

// ---------------------------------------------------------------------------
// Several forward steps per launch, full-width workgroups
// ---------------------------------------------------------------------------
//
// For grids up to 1024 columns one workgroup holds a whole row: wavefront w
// covers columns 64 w - off .. 64 w - off + 63 (off = 1 when nx = 64 k + 1,
// so that column nx - 1 never sits on lane 0 and its Mur update finds its
// west neighbour in the same wavefront; otherwise 0 and rows are 256-byte
// aligned). KL time levels are pipelined down a row segment of one shot:
// at iteration y, level j (1..KL) computes row y - j + 1 from the rows of
// level j - 1 (level 0 = the input field u, level -1 = u_old) that are
// already in registers, so the inputs are read once and only the levels
// asked for are written. The wavefronts march in lockstep, one barrier per
// iteration, and pass the values next to their block edges (the newest row
// of every level but the last) to their neighbours through LDS,
// double-buffered by iteration parity: nothing is loaded or recomputed
// twice across columns. Rows within KL - 1 of a segment's ends are
// recomputed by the neighbouring segment. Every cell gets the arithmetic of
// the single-step kernel, so the results are the same bits.
//
// Stores are raw buffer stores; lanes outside the grid use an offset past
// the buffer's end, which the hardware drops (planes are < 2 GiB).

constexpr int kMaxWideWaves = 16;
constexpr int kMaxLevels = 4;
constexpr int kDropOffset = (int)0x80000000u;

// Edge values of one wavefront side: the newest row of each level, then (for
// Laplacian histories) the row before it.
struct alignas(16) Edge4 {
    float v[2 * kMaxLevels];
};
// [iteration parity][wavefront slot: 0 and nw + 1 are zero pads][west, east]
typedef Edge4 EdgeRow[kMaxWideWaves + 2][2];

// A prefetched slot value copied into a register of its own (early clobber):
// the slot's register is then free for the refill, so a march keeps each slot
// in one register across its loop and the latch needs no copy of a value the
// refill just loaded (such a copy waits there for that load and, the counter
// being shared, for every store before it).
__device__ __forceinline__ float take(float v)
{
    float r;
    asm volatile("v_mov_b32 %0, %1" : "=&v"(r) : "v"(v));
    return r;
}

__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, float v, int voff, int soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, soff, 0);
}

// Rows loaded ahead of the march; three, the period of the row rotation,
// so that an unrolled group needs no register moves.
constexpr int kPfW = 3;

// Which levels a launch stores: kStoreAny = those with a non-NULL output
// (checked per level), kStoreLastTwo = the last two, kStoreAll = every level,
// kStoreFootprint = the last two everywhere and the others (kept only for the
// receiver gather after the launch) in the rows and wavefronts p.fp_rows
// marks.
enum : int { kStoreAny = 0, kStoreLastTwo = 1, kStoreAll = 2, kStoreFootprint = 3 };

// XE: the wavefront holds column 0 or nx - 1 (Mur side updates); YE: the
// segment's levels reach rows 0 / 1 or ny - 1 (Mur row updates).
template <int KL, int SM, bool XE, bool YE, unsigned F>
__device__ __forceinline__ void fwdm_wave(const StepParams &p, int lane, int wave, int grp,
                                          int seg, EdgeRow *ex)
{
    constexpr bool LAP = (F & kLap) != 0;
    constexpr bool BND = (F & kBnd) != 0;
    const int W = p.nx, H = p.ny;

    const int x = wave * kLanes - p.wide_off + lane;
    // only the first and last wavefronts of a row (XE) reach past the grid
    const bool owned = !XE || ((x >= 0) & (x < W));
    // Laplacian histories need values outside the grid to be zero: lanes
    // beyond the side columns and rows beyond the top and bottom (they only
    // feed boundary cells, which the Mur updates replace, so the fields do
    // not change)
    auto in_grid = [&](int row) -> bool {
        return (!XE || owned) && (!YE || ((row >= 0) & (row < H)));
    };
    const int xo = min(max(x, 0), W - 1) * 4;
    const int so = owned ? x * 4 : kDropOffset;
    int r0, r1;
    seg_rows(p, seg, r0, r1);
    const int y_begin = max(r0 - (KL - 1), 0);
    const int y_end = r1 + KL - 2;  // level KL reaches row r1 - 1
    const size_t plane = (size_t)H * W;
    const unsigned plane_bytes = (unsigned)(plane * sizeof(float));
    const int row_bytes = W * 4;
    auto row_off = [&](int yy) -> int { return min(max(yy, 0), H - 1) * row_bytes; };
    const __amdgpu_buffer_rsrc_t rk = make_rsrc(p.c2dt2, plane_bytes);
    const float *mur = p.mur;
    const __amdgpu_buffer_rsrc_t rm = make_rsrc(p.mur, (unsigned)(2 * H * sizeof(float)));
    const int mo = (x == 0) ? 0 : H * 4;
    auto mrow = [&](int yy) -> int { return min(max(yy, 0), H - 1) * 4; };
    float kb_x = 0.0f, kt_x = 0.0f;
    if constexpr (YE) {
        kb_x = mur[2 * H + xo / 4];
        kt_x = mur[2 * H + W + xo / 4];
    }
    // LDS slots the neighbour values come from: lane 0 reads the east edge
    // of the wavefront to the west, lane 63 the west edge of the one to the
    // east (pads beyond the grid); other lanes read their own wavefront's
    // slots and ignore the values (distinct addresses for the two reads).
    const int w_slot = lane == 0 ? wave : wave + 1;
    const int w_side = lane == 0 ? 1 : 0;
    const int e_slot = lane == kLanes - 1 ? wave + 2 : wave + 1;
    const int e_side = lane == kLanes - 1 ? 0 : 1;
    const bool edge_lane = (lane == 0) | (lane == kLanes - 1);
    const int my_side = lane == kLanes - 1 ? 1 : 0;
    const int s_begin = grp * p.shots_per_group;
    const int s_end = min(p.n_shots, s_begin + p.shots_per_group);
    for (int s = s_begin; s < s_end; ++s) {
        const size_t soff = (size_t)s * plane;
        const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.f_old + soff, plane_bytes);
        const __amdgpu_buffer_rsrc_t rb = make_rsrc(p.f_cur + soff, plane_bytes);
        __amdgpu_buffer_rsrc_t ro[KL], ro2[KL], rl[KL];
#pragma unroll
        for (int j = 0; j < KL; ++j) {
            // integer arithmetic: the pointer may be NULL (level not stored)
            ro[j] = make_rsrc((const float *)((uintptr_t)p.lv_out[j] + soff * sizeof(float)),
                              plane_bytes);
            ro2[j] = make_rsrc((const float *)((uintptr_t)p.lv_out2[j] + soff * sizeof(float)),
                               plane_bytes);
            rl[j] = make_rsrc((const float *)((uintptr_t)p.lv_lap[j] + soff * sizeof(float)),
                              plane_bytes);
        }
        const int sr = p.src_rows[s], sc = p.src_cols[s];
        const int src_extra = transmitter_rows(p);
        float sv[KL];
#pragma unroll
        for (int j = 0; j < KL; ++j) sv[j] = p.src_values[p.src_index + j];
        // lv[j][0..2]: level j's newest row and the two before it; level 0
        // is the input field (its newest row at iteration y is y + 1)
        float lv[KL + 1][3];
        float kh[KL], mh[KL];  // coefficient rows y, y - 1, ...
#pragma unroll
        for (int j = 0; j <= KL; ++j) lv[j][0] = lv[j][1] = lv[j][2] = 0.0f;
#pragma unroll
        for (int j = 0; j < KL; ++j) kh[j] = mh[j] = 0.0f;
        lv[0][0] = bload(rb, xo, row_off(y_begin));
        lv[0][1] = bload(rb, xo, row_off(y_begin - 1));
        if constexpr (LAP) {
            lv[0][0] = in_grid(y_begin) ? lv[0][0] : 0.0f;
            lv[0][1] = in_grid(y_begin - 1) ? lv[0][1] : 0.0f;
        }
        float qb[kPfW], qa[kPfW], qk[kPfW], qm[kPfW];
#pragma unroll
        for (int i = 0; i < kPfW; ++i) {
            __builtin_amdgcn_sched_barrier(0);
            qb[i] = bload(rb, xo, row_off(y_begin + 1 + i));
            qa[i] = bload(ra, xo, row_off(y_begin + i));
            qk[i] = bload(rk, xo, row_off(y_begin + i));
            qm[i] = 0.0f;
            if constexpr (XE) qm[i] = bload(rm, mo, mrow(y_begin + i));
        }
        __builtin_amdgcn_sched_barrier(0);
        if (edge_lane) {
            Edge4 e;
#pragma unroll
            for (int j = 0; j < 2 * kMaxLevels; ++j) e.v[j] = 0.0f;
            e.v[0] = lv[0][0];
            e.v[kMaxLevels] = lv[0][1];
            ex[y_begin & 1][wave + 1][my_side] = e;
        }
        __syncthreads();
        auto store_row = [&](int j, int rho, float v) {
            bool stored = SM == kStoreAll || (SM == kStoreLastTwo && j >= KL - 2) ||
                          (SM == kStoreAny && p.lv_out[j] != nullptr);
            if constexpr (SM == kStoreFootprint) {
                // row and wavefront are uniform: a scalar load of the row's mask
                stored = j >= KL - 2 || (((unsigned)((const kci *)p.fp_rows)[rho] >> wave) & 1u) != 0;
                if constexpr (BND) {
                    // the boundary cells too (the ring is taken from the fields):
                    // rows 0 and H - 1 whole, columns 0 and W - 1 by their lanes
                    if (!stored && YE && ((rho == 0) | (rho == H - 1))) stored = true;
                    if (!stored && XE) {
                        bstore(ro[j], v, ((x == 0) | (x == W - 1)) ? so : kDropOffset,
                               rho * row_bytes);
                        return;
                    }
                }
            }
            if (stored) bstore(ro[j], v, so, rho * row_bytes);
            if constexpr ((F & kOut2) != 0)
                if (p.lv_out2[j] != nullptr) bstore(ro2[j], v, so, rho * row_bytes);
        };
        auto iteration = [&](int y, float &sb, float &sa, float &sk, float &sm) {
#pragma unroll
            for (int j = 0; j <= KL; ++j) {
                lv[j][2] = lv[j][1];
                lv[j][1] = lv[j][0];
            }
#pragma unroll
            for (int j = KL - 1; j >= 1; --j) {
                kh[j] = kh[j - 1];
                mh[j] = mh[j - 1];
            }
            float b1 = sb, a0 = sa, k0 = sk, m0 = sm;
            asm volatile("" : "+v"(b1), "+v"(a0), "+v"(k0), "+v"(m0));
            __builtin_amdgcn_sched_barrier(0);
            lv[0][0] = LAP && !in_grid(y + 1) ? 0.0f : b1;
            kh[0] = k0;
            mh[0] = m0;
            const int yf = min(y + kPfW, y_end);
            sb = bload(rb, xo, row_off(yf + 1));
            sa = bload(ra, xo, row_off(yf));
            sk = bload(rk, xo, row_off(yf));
            if constexpr (XE) sm = bload(rm, mo, mrow(yf));
            const Edge4 *hw = &ex[y & 1][w_slot][w_side];
            const Edge4 *he = &ex[y & 1][e_slot][e_side];
#pragma unroll
            for (int j = 1; j <= KL; ++j) {
                const int rho = y - j + 1;
                const float u = lv[j - 1][1], un = lv[j - 1][0], us = lv[j - 1][2];
                const float old = j == 1 ? a0 : lv[j - 2][2];
                // west and east neighbour values read per level into registers
                // that the lane shifts overwrite in place
                const float uw = from_west(u, hw->v[j - 1]), ue = from_east(u, he->v[j - 1]);
                const float lp = lap5(uw, ue, us, un, u, p.idx2, p.idy2, p.cen);
                float v = __fmaf_rn(kh[j - 1], lp, leap(u, old));
                if (__builtin_expect((unsigned)(rho - sr) <= (unsigned)src_extra, 0))
                    v = add_transmitter(p, s, v, rho, x, sr, sc, sv[j - 1]);
                if constexpr (XE) {
                    // rows of segments away from the top and bottom are interior
                    if (!YE || (rho >= 1 && rho <= H - 2)) v = mur_sides(v, u, x, W, mh[j - 1]);
                }
                if constexpr (YE) {
                    if (rho == H - 1) v = mur_row(mur, W, H, x, kt_x, u, us, lv[j][1], 2);
                    if (rho == 1) {
                        float v0 = mur_row(mur, W, H, x, kb_x, us, u, v, 0);
                        if constexpr (LAP) v0 = owned ? v0 : 0.0f;
                        lv[j][1] = v0;
                        if (r0 == 0) store_row(j - 1, 0, v0);
                        if constexpr (LAP) {
                            // row 0 of the input is final only now (and its edge
                            // values came as the previous row): its Laplacian
                            const float z = us;  // input row 0
                            const float zw = from_west(z, hw->v[kMaxLevels + j - 1]);
                            const float ze = from_east(z, he->v[kMaxLevels + j - 1]);
                            const float l0 = lap5(zw, ze, 0.0f, u, z, p.idx2, p.idy2, p.cen);
                            if ((r0 == 0) & (p.lv_lap[j - 1] != nullptr)) bstore(rl[j - 1], l0, so, 0);
                        }
                    }
                }
                if constexpr (LAP) v = in_grid(rho) ? v : 0.0f;
                lv[j][0] = v;
                if ((rho >= r0) & (rho < r1) & (rho > 0)) store_row(j - 1, rho, v);
                if constexpr (LAP) {
                    if ((p.lv_lap[j - 1] != nullptr) & (rho >= r0) & (rho < r1) & (rho > 0))
                        bstore(rl[j - 1], lp, so, rho * row_bytes);
                }
            }
            if (edge_lane) {
                Edge4 e;
#pragma unroll
                for (int j = 0; j < kMaxLevels; ++j) e.v[j] = j < KL ? lv[j][0] : 0.0f;
                if constexpr (LAP && YE) {
#pragma unroll
                    for (int j = 0; j < kMaxLevels; ++j)
                        e.v[kMaxLevels + j] = j < KL ? lv[j][1] : 0.0f;
                }
                ex[(y + 1) & 1][wave + 1][my_side] = e;
            }
            __syncthreads();
        };
        // exactly y_end - y_begin + 1 iterations (one barrier each), as the
        // lean march runs
        static_assert(kPfW == 3, "the march below takes its prefetch slots in threes");
        for (int y = y_begin;; y += kPfW) {
            iteration(y, qb[0], qa[0], qk[0], qm[0]);
            if (y + 1 > y_end) break;
            iteration(y + 1, qb[1], qa[1], qk[1], qm[1]);
            if (y + 2 > y_end) break;
            iteration(y + 2, qb[2], qa[2], qk[2], qm[2]);
            if (y + 3 > y_end) break;
        }
    }
}

// Lean forward march: footprint stores (the last two levels everywhere, the
// others in the rows and wavefronts holding receiver footprint cells, and with
// BND on the boundary columns too), in a segment whose march reaches neither
// row 1 nor row H - 1, of a row spanning at least two wavefronts. SIDE 0 = a
// wavefront holding neither column 0 nor W - 1, 1 = the one holding column 0,
// 2 = the one holding W - 1. Per cell it does fwdm_wave's arithmetic in the
// same order (the same bits); the bookkeeping is that of adjr_lean: rings of
// registers made static by a six-iteration unroll (no latch copies), the
// prefetch loading straight into the rings, a compile-time prologue for the
// rows below the segment and no row tests after it (the last iterations'
// rows beyond the segment are the next segment's, made from the same inputs:
// the same bits), the footprint mask of the next row loaded one iteration
// ahead into a scalar register (the barrier's wait covers it), every store
// issued with a dropped lane offset where nothing is stored (exact wait
// counts), and a one-sided Mur update on the side wavefronts.
template <int KL, int SIDE, unsigned F>
__device__ __forceinline__ void fwdm_lean(const StepParams &p, int lane, int wave, int grp,
                                          int seg, EdgeRow *ex)
{
    constexpr bool BND = (F & kBnd) != 0;
    constexpr int kR = 6;             // ring length = iterations per unrolled group
    constexpr int kPro = 2 * KL - 2;  // iterations before every stored row is owned
    constexpr int kPf = SIDE == 0 ? 3 : 2;  // rows loaded ahead (side marches hold more)
    static_assert(kR % kPf == 0 && kPf + 3 <= kR, "prefetched rows must fit the rings");
    const int W = p.nx, H = p.ny;
    const int x = wave * kLanes - p.wide_off + lane;
    const bool owned = SIDE == 0 || ((x >= 0) & (x < W));
    const int xo = (SIDE == 0 ? x : min(max(x, 0), W - 1)) * 4;
    const int so = owned ? x * 4 : kDropOffset;
    const bool side_lane = SIDE == 1 ? x == 0 : (SIDE == 2 ? x == W - 1 : false);
    const int mo = SIDE == 2 ? H * 4 : 0;
    int r0, r1;
    seg_rows(p, seg, r0, r1);
    const int y_begin = r0 - (KL - 1);
    const int y_end = r1 + KL - 2;  // level KL reaches row r1 - 1
    const int n_it = y_end - y_begin + 1;
    const size_t plane = (size_t)H * W;
    const unsigned plane_bytes = (unsigned)(plane * sizeof(float));
    const int row_bytes = W * 4;
    const __amdgpu_buffer_rsrc_t rk = make_rsrc(p.c2dt2, plane_bytes);
    const __amdgpu_buffer_rsrc_t rm = make_rsrc(p.mur, (unsigned)(2 * H * sizeof(float)));
    const kci *fp = (const kci *)p.fp_rows;
    const float idx2 = p.idx2, idy2 = p.idy2, cen = p.cen;
    const int w_slot = lane == 0 ? wave : wave + 1;
    const int w_side = lane == 0 ? 1 : 0;
    const int e_slot = lane == kLanes - 1 ? wave + 2 : wave + 1;
    const int e_side = lane == kLanes - 1 ? 0 : 1;
    const bool edge_lane = (lane == 0) | (lane == kLanes - 1);
    const int my_side = lane == kLanes - 1 ? 1 : 0;
    // edge rows by iteration parity: iteration c reads ex_c[c & 1] and writes
    // ex_c[(c + 1) & 1] (the rings' length is even, so c & 1 is static)
    EdgeRow *const ex_c[2] = {ex + (y_begin & 1), ex + ((y_begin + 1) & 1)};
    const int s_begin = grp * p.shots_per_group;
    const int s_end = min(p.n_shots, s_begin + p.shots_per_group);
    const int src_extra = transmitter_rows(p);
    for (int s = s_begin; s < s_end; ++s) {
        const size_t soff = (size_t)s * plane;
        const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.f_old + soff, plane_bytes);
        const __amdgpu_buffer_rsrc_t rb = make_rsrc(p.f_cur + soff, plane_bytes);
        __amdgpu_buffer_rsrc_t ro[KL];
#pragma unroll
        for (int j = 0; j < KL; ++j)
            ro[j] = make_rsrc((const float *)((uintptr_t)p.lv_out[j] + soff * sizeof(float)),
                              plane_bytes);
        const int sr = p.src_rows[s], sc = p.src_cols[s];
        float sv[KL];
#pragma unroll
        for (int j = 0; j < KL; ++j) sv[j] = p.src_values[p.src_index + j];
        // ring slot of iteration c = c mod kR: u[0] = the input's row y + 1,
        // ua = the older input's row y, kh / mh = dt^2 c^2 and the Mur side
        // coefficient of row y, fm = the footprint mask of row y; u[j] (j >= 1)
        // = level j's row made at iteration c (row y - j + 1). The prefetch
        // loads u[0] and ua straight into their rings; kh and mh stay in use
        // for KL iterations and go through slots qk, qm.
        float u[KL + 1][kR], ua[kR], kh[kR], mh[kR], qk[kPf], qm[kPf];
        int fm[kR];
#pragma unroll
        for (int k = 0; k < kR; ++k) {
#pragma unroll
            for (int j = 0; j <= KL; ++j) u[j][k] = 0.0f;
            ua[k] = kh[k] = mh[k] = 0.0f;
            fm[k] = 0;
        }
        // iterations -1 and -2: rows y_begin and y_begin - 1
        u[0][kR - 1] = bload(rb, xo, y_begin * row_bytes);
        u[0][kR - 2] = bload(rb, xo, (y_begin - 1) * row_bytes);
#pragma unroll
        for (int i = 0; i < kPf; ++i) {
            __builtin_amdgcn_sched_barrier(0);
            const int o0 = (y_begin + i) * row_bytes;
            u[0][i] = bload(rb, xo, o0 + row_bytes);
            ua[i] = bload(ra, xo, o0);
            qk[i] = bload(rk, xo, o0);
            qm[i] = 0.0f;
            if constexpr (SIDE != 0) qm[i] = bload(rm, mo, (y_begin + i) * 4);
        }
        __builtin_amdgcn_sched_barrier(0);
        fm[0] = fp[y_begin];
        if (edge_lane) {
            Edge4 e;
#pragma unroll
            for (int j = 0; j < 2 * kMaxLevels; ++j) e.v[j] = 0.0f;
            e.v[0] = u[0][kR - 1];
            e.v[kMaxLevels] = u[0][kR - 2];
            (*ex_c[0])[wave + 1][my_side] = e;
        }
        __syncthreads();
        // iteration c with ring slot cm = c mod kR; pc = c in the first kPro
        // iterations (row tests at compile time), -1 afterwards (no tests)
        auto iteration = [&](int cm, int pc, int c) __attribute__((always_inline)) {
            const int y = y_begin + c;
            const int cn = cm, cc = (cm + kR - 1) % kR, cs = (cm + kR - 2) % kR;
            const int q = cm % kPf;
            const float k_new = take(qk[q]);
            const float m_new = SIDE != 0 ? take(qm[q]) : 0.0f;
            __builtin_amdgcn_sched_barrier(0);
            {
                const int f = (cm + kPf) % kR;
                const int yp = min(y + kPf, y_end);
                const int o0 = yp * row_bytes;
                u[0][f] = bload(rb, xo, o0 + row_bytes);
                ua[f] = bload(ra, xo, o0);
                qk[q] = bload(rk, xo, o0);
                if constexpr (SIDE != 0) qm[q] = bload(rm, mo, yp * 4);
            }
            __builtin_amdgcn_sched_barrier(0);
            kh[cm] = k_new;
            mh[cm] = m_new;
            // the next iteration's footprint row (scalar; the barrier at the
            // end of this iteration waits for it)
            fm[(cm + 1) % kR] = fp[min(y + 1, y_end)];
            const bool near_src = (unsigned)(y - sr) <= (unsigned)(KL - 1 + src_extra);
            const EdgeRow &er = *ex_c[cm & 1];
#pragma unroll
            for (int j = 1; j <= KL; ++j) {
                const int rho = y - j + 1;
                const int kj = (cm + kR - (j - 1)) % kR;  // row rho's coefficients and mask
                const float uc = u[j - 1][cc], un = u[j - 1][cn], us = u[j - 1][cs];
                const float old = j == 1 ? ua[cm] : u[j - 2][cs];
                const float uw = from_west(uc, er[w_slot][w_side].v[j - 1]);
                const float ue = from_east(uc, er[e_slot][e_side].v[j - 1]);
                const float lp = lap5(uw, ue, us, un, uc, idx2, idy2, cen);
                float v = __fmaf_rn(kh[kj], lp, leap(uc, old));
                if (__builtin_expect(near_src, 0) && (unsigned)(rho - sr) <= (unsigned)src_extra)
                    v = add_transmitter(p, s, v, rho, x, sr, sc, sv[j - 1]);
                if constexpr (SIDE == 1) {  // column 0: from column 1's new and current values
                    const float ve = from_east(v, 0.0f), ce = from_east(uc, 0.0f);
                    v = side_lane ? ce + mh[kj] * (ve - uc) : v;
                } else if constexpr (SIDE == 2) {  // column W - 1: from column W - 2's
                    const float vw = from_west(v, 0.0f), cw = from_west(uc, 0.0f);
                    v = side_lane ? cw + mh[kj] * (vw - uc) : v;
                }
                u[j][cm] = v;
                if (pc < 0 || pc >= KL + j - 2) {  // row rho >= r0
                    int off = so;
                    if (j < KL - 1) {  // footprint rows (and the ring's columns)
                        const bool fp_here = ((unsigned)fm[kj] >> wave) & 1u;
                        off = (fp_here || (BND && side_lane)) ? so : kDropOffset;
                    }
                    bstore(ro[j - 1], v, off, rho * row_bytes);
                }
            }
            if (edge_lane) {  // the newest row of each level (the rest is not read)
                float e[kMaxLevels];
#pragma unroll
                for (int j = 0; j < kMaxLevels; ++j) e[j] = j < KL ? u[j][cm] : 0.0f;
                *reinterpret_cast<float4 *>(&(*ex_c[(cm + 1) & 1])[wave + 1][my_side].v[0]) =
                    make_float4(e[0], e[1], e[2], e[3]);
            }
            __syncthreads();
        };
#pragma unroll
        for (int c = 0; c < kPro; ++c) iteration(c % kR, c, c);
        // the rest in groups of kR from ring slot kPro mod kR; exactly n_it
        // iterations in all (one barrier each, as in fwdm_wave)
        int c = kPro;
        for (;;) {
#pragma unroll
            for (int k = 0; k < kR; ++k) {
                iteration((kPro + k) % kR, -1, c);
                if (++c == n_it) goto march_done;
            }
        }
    march_done:;
    }
}

// Workgroup b of a full-width multi-step forward launch of one problem.
template <int KL, int SM, unsigned F>
__device__ __forceinline__ void fwdm_block(const StepParams &p, int b, EdgeRow *ex)
{
    if constexpr ((F & kGather) != 0) {
        if (b >= p.n_stencil_blocks) {
            gather_pairs(p, (b - p.n_stencil_blocks) * (int)blockDim.x + (int)threadIdx.x);
            return;
        }
    }
    const int wblk = (b % kXcds) * p.chunk + b / kXcds;
    if (wblk >= p.n_work) return;  // uniform over the workgroup
    const int t = (int)threadIdx.x;
    if (t < 2 * 2 * 2 * kMaxLevels) {  // zero pads at both ends, both parities
        const int par = t / (4 * kMaxLevels), side = (t / (2 * kMaxLevels)) & 1,
                  j = t % (2 * kMaxLevels);
        ex[par][0][side].v[j] = 0.0f;
        ex[par][p.wide_waves + 1][side].v[j] = 0.0f;
    }
    const int lane = t & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane(t / kLanes);
    // shot group fastest: consecutive workgroups share the coefficient rows
    const int grp = wblk % p.n_groups;
    const int seg = wblk / p.n_groups;
    int r0, r1;
    seg_rows(p, seg, r0, r1);
    const bool xe = (wave == 0) || (wave == p.wide_waves - 1);
    const bool ye = (r0 <= KL) || (r1 + KL - 2 >= p.ny - 1);
    if constexpr (SM == kStoreFootprint && (F & (kLap | kOut2)) == 0) {
        // lean marches away from rows 1 / H - 1 (the same bits)
        if (!ye && p.fwdm_lean) {
            if (wave == 0) fwdm_lean<KL, 1, F>(p, lane, wave, grp, seg, ex);
            else if (wave == p.wide_waves - 1) fwdm_lean<KL, 2, F>(p, lane, wave, grp, seg, ex);
            else fwdm_lean<KL, 0, F>(p, lane, wave, grp, seg, ex);
            return;
        }
    }
    if (xe) {
        if (ye) fwdm_wave<KL, SM, true, true, F>(p, lane, wave, grp, seg, ex);
        else fwdm_wave<KL, SM, true, false, F>(p, lane, wave, grp, seg, ex);
    } else {
        if (ye) fwdm_wave<KL, SM, false, true, F>(p, lane, wave, grp, seg, ex);
        else fwdm_wave<KL, SM, false, false, F>(p, lane, wave, grp, seg, ex);
    }
}

// At most 64 VGPRs (eight wavefronts per SIMD): two full-width workgroups
// per CU, so one marches while the other waits at its barrier.
template <int KL, int SM, unsigned F>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8)))
fwdm_kernel(StepParams p)
{
    __shared__ EdgeRow ex[2];
    fwdm_block<KL, SM, F>(p, (int)blockIdx.x, ex);
}

// Batched full-width launch: entry e of `table` is a problem's launch
// (all entries have the same row width, so the same workgroup size).
template <int KL, int SM, unsigned F>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8)))
fwdm_batch_kernel(const StepParams *table, const int *first, int n)
{
    __shared__ EdgeRow ex[2];
    const int b = (int)blockIdx.x;
    const int e = __builtin_amdgcn_readfirstlane(batch_entry(first, n, b));
    fwdm_block<KL, SM, F>(batch_params(table, e), b - first[e], ex);
}

// Workgroup b of a two-step forward launch of one problem.
template <unsigned F>
__device__ __forceinline__ void fwd2_block(const StepParams &p, int b)
{
    if constexpr ((F & kGather) != 0) {
        if (b >= p.n_stencil_blocks) {
            gather_pairs(p, (b - p.n_stencil_blocks) * (int)blockDim.x + (int)threadIdx.x);
            return;
        }
    }
    const int wblk = (b % kXcds) * p.chunk + b / kXcds;
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kLanes));
    const int item = wblk * kWavesPerBlock + wave;
    if (item >= p.n_work) return;
    // shot group fastest: the waves of a workgroup share coefficient rows
    const int grp = item % p.n_groups;
    const int t = item / p.n_groups;
    const int cb = t % p.n_cb;
    const int seg = t / p.n_cb;
    const int xs = min(cb * kLanes, max(p.nx - kLanes, 0));
    int r0, r1;
    seg_rows(p, seg, r0, r1);
    // Mur work: column 0 or W-1 among the block and halo columns, or the
    // first / last row among the segment's rows and its halo rows
    const bool edge = (xs <= 1) || (xs + kLanes >= p.nx - 1) || (r0 == 0) || (r1 >= p.ny - 1);
    if (edge)
        fwd2_wave<true, F>(p, lane, grp, cb, seg);
    else
        fwd2_wave<false, F>(p, lane, grp, cb, seg);
}

template <unsigned F>
__global__ void __launch_bounds__(256) fwd2_kernel(StepParams p)
{
    fwd2_block<F>(p, (int)blockIdx.x);
}

// One launch advancing several problems by two forward steps each: entry e
// of `table` holds problem e's launch parameters (as the single-problem
// kernel receives them).
template <unsigned F>
__global__ void __launch_bounds__(256) fwd2_batch_kernel(const StepParams *table, const int *first,
                                                         int n)
{
    const int b = (int)blockIdx.x;
    const int e = __builtin_amdgcn_readfirstlane(batch_entry(first, n, b));
    fwd2_block<F>(batch_params(table, e), b - first[e]);
}

// ---------------------------------------------------------------------------
// Two adjoint steps per launch (temporal blocking)
// ---------------------------------------------------------------------------
//
// Steps n (first) and n-1 (second) of the reverse run, from A = lam before
// step n (older) and B = lam after it... that is, the state (A, B) of the
// reverse run: C = 2B - A + dt^2 Lap(c^2 B) + residual_n, D = 2C - B +
// dt^2 Lap(c^2 C) + residual_{n-1}, each followed by the forward Mur update.
// Same 64-column marches as the two-step forward kernel; Lap(c^2 .) works on
// products c^2 * lam kept per row, so the halo columns also need c^2 beside
// them (loaded by lanes 1 / 62 like B). Bilinear residual injection goes
// through the scatter table in the single-step kernel's order, on the block
// and halo columns. Per cell and step the arithmetic equals the single-step
// kernel's, so C and D are bit-identical to two single steps.
//
// Imaging: C is imaged against F1 (forward field after step n-1) and D
// against F2 (after step n-2), as g += scale sum_s (C Lap F1 + D Lap F2).
// The four wavefronts of a workgroup take disjoint shots of one shot group at
// the same rows and columns; each adds its shots' terms per cell into its own
// LDS rows, and at the end the workgroup adds the four in a fixed order into
// the group's partial image (at most kMaxImageGroups), so the gradient is
// deterministic. A cell's C term is added one row after C is final and its D
// term one row after that (row 0 of each step is finalised a row late).

template <bool EDGE, bool IMG>
__device__ __forceinline__ void adj2_wave(const StepParams &p, int lane, int wave, int grp,
                                          int cb, int seg, float *acc)
{
    const int W = p.nx, H = p.ny;
    const int own0 = cb * kLanes;
    const int xs = min(own0, max(W - kLanes, 0));
    const int x = xs + lane;
    const bool xin = x < W;
    const bool owned = xin & (x >= own0);
    const int xo = min(x, W - 1) * 4;
    const bool hl = lane == 0, hr = lane == kLanes - 1;
    const int xh = hl ? xs - 1 : (hr ? xs + kLanes : x);
    const int xq = lane == 1 ? xs - 2 : (lane == kLanes - 2 ? xs + kLanes + 1 : xh);
    const bool xh_in = (xh >= 0) & (xh < W);
    const bool xq_in = (xq >= 0) & (xq < W);
    const int ho = min(max(xh, 0), W - 1) * 4;
    const int qo = min(max(xq, 0), W - 1) * 4;
    int r0, r1;
    seg_rows(p, seg, r0, r1);
    const int y_start = r0 == 0 ? 0 : r0 - 1;
    const int y_end = r1 + 1;  // the D term of row r1 - 1 is imaged at row r1 + 1
    const size_t plane = (size_t)H * W;
    const unsigned plane_bytes = (unsigned)(plane * sizeof(float));
    auto row_off = [&](int yy) -> int { return min(max(yy, 0), H - 1) * W * 4; };
    const __amdgpu_buffer_rsrc_t rk = make_rsrc(p.c2, plane_bytes);
    const __amdgpu_buffer_rsrc_t re = make_rsrc((const float *)p.cell_map, plane_bytes);
    const float *mur = p.mur;
    const __amdgpu_buffer_rsrc_t rm = make_rsrc(p.mur, (unsigned)(2 * H * sizeof(float)));
    const int mo = (x == 0) ? 0 : H * 4;
    auto mrow = [&](int yy) -> int { return min(max(yy, 0), H - 1) * 4; };
    const float dt2 = p.dt2;
    float kb_x = 0.0f, kt_x = 0.0f, kb_h = 0.0f, kt_h = 0.0f;
    float kx1 = 0.0f, ky1 = 0.0f, kx3 = 0.0f, ky3 = 0.0f;
    const bool h_right_edge = hr & (xh == W - 1);
    if constexpr (EDGE) {
        kb_x = mur[2 * H + xo / 4];
        kt_x = mur[2 * H + W + xo / 4];
        kb_h = mur[2 * H + ho / 4];
        kt_h = mur[2 * H + W + ho / 4];
        const float *kx = mur + 2 * H + 2 * W;
        kx1 = kx[1];
        ky1 = kx[4 + 1];
        kx3 = kx[3];
        ky3 = kx[4 + 3];
    }
    // does any block or halo cell of these rows receive a residual?
    bool inject;
    {
        bool mine = false;
        for (int yy = y_start; yy <= r1 && yy < H; ++yy) {
            const int e = __float_as_int(bload(re, xo, row_off(yy)));
            const int eh = __float_as_int(bload(re, ho, row_off(yy)));
            mine |= (xin & (e >= 0)) | (xh_in & (eh >= 0));
        }
        inject = __any(mine);
    }
    const int n_rows = r1 - r0;
    if constexpr (IMG) {
        for (int i = 0; i < n_rows; ++i) acc[i * kLanes + lane] = 0.0f;
    }
    const float *img1 = IMG ? p.img : nullptr;
    const float *img2 = IMG ? p.img_b : nullptr;
    const int s_begin = grp * p.shots_per_group + wave * p.shots_per_wave;
    const int s_end = min(min(p.n_shots, s_begin + p.shots_per_wave),
                          (grp + 1) * p.shots_per_group);
    for (int s = s_begin; s < s_end; ++s) {
        const size_t soff = (size_t)s * plane;
        const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.f_old + soff, plane_bytes);
        const __amdgpu_buffer_rsrc_t rb = make_rsrc(p.f_cur + soff, plane_bytes);
        const __amdgpu_buffer_rsrc_t rf1 = make_rsrc(img1 != nullptr ? img1 + soff : p.f_cur,
                                                     plane_bytes);
        const __amdgpu_buffer_rsrc_t rf2 = make_rsrc(img2 != nullptr ? img2 + soff : p.f_cur,
                                                     plane_bytes);
        const bool use1 = img1 != nullptr, use2 = img2 != nullptr;
        gf *outc = gstore_ptr(p.f_out) + soff;
        gf *outd = gstore_ptr(p.f_out_b) + soff;
        const float *res1 = p.resid + (size_t)s * p.n_rec;
        const float *res2 = p.resid_b + (size_t)s * p.n_rec;
        // residual injected at a cell: sum over its scatter entries, in order
        auto injected = [&](const float *res, int e) -> float {
            float a = 0.0f;
            for (int k = p.entry_start[e]; k < p.entry_start[e + 1]; ++k)
                a += res[p.pair_rec[k]] * p.pair_w[k];
            return a;
        };
        float bm, b0, qm1, q0, km, k0, kqm1, kq0;
        float f1m2 = 0.0f, f1m1 = 0.0f, f1hm1 = 0.0f, f2m3 = 0.0f, f2m2 = 0.0f, f2hm2 = 0.0f;
        {
            const float t0 = bload(rb, xo, row_off(y_start - 1));
            const float t1 = bload(rb, xo, row_off(y_start));
            const float t2 = bload(rb, qo, row_off(y_start - 1));
            const float t3 = bload(rb, qo, row_off(y_start));
            const float u0 = bload(rk, xo, row_off(y_start - 1));
            const float u1 = bload(rk, xo, row_off(y_start));
            const float u2 = bload(rk, qo, row_off(y_start - 1));
            const float u3 = bload(rk, qo, row_off(y_start));
            const bool up = y_start >= 1;
            bm = (xin & up) ? t0 : 0.0f;
            b0 = xin ? t1 : 0.0f;
            qm1 = (xq_in & up) ? t2 : 0.0f;
            q0 = xq_in ? t3 : 0.0f;
            km = (xin & up) ? u0 : 0.0f;
            k0 = xin ? u1 : 0.0f;
            kqm1 = (xq_in & up) ? u2 : 0.0f;
            kq0 = xq_in ? u3 : 0.0f;
            if constexpr (IMG) {
                const float v0 = bload(rf1, xo, row_off(y_start - 2));
                const float v1 = bload(rf1, xo, row_off(y_start - 1));
                const float v2 = bload(rf1, ho, row_off(y_start - 1));
                const float w0 = bload(rf2, xo, row_off(y_start - 3));
                const float w1 = bload(rf2, xo, row_off(y_start - 2));
                const float w2 = bload(rf2, ho, row_off(y_start - 2));
                f1m2 = (xin & (y_start >= 2)) ? v0 : 0.0f;
                f1m1 = (xin & up) ? v1 : 0.0f;
                f1hm1 = (xh_in & up) ? v2 : 0.0f;
                f2m3 = (xin & (y_start >= 3)) ? w0 : 0.0f;
                f2m2 = (xin & (y_start >= 2)) ? w1 : 0.0f;
                f2hm2 = (xh_in & (y_start >= 2)) ? w2 : 0.0f;
            }
        }
        // Slots, kPf rows ahead, consumed and refilled in place (as in the
        // forward kernel): B, c^2 at row y+1 (block and halo/outer columns);
        // A and the scatter-table entry at row y (block and halo columns);
        // F1 at row y and F2 at row y-1 (block and halo columns).
        float qb[kPf], qq[kPf], qk[kPf], qkq[kPf], qa[kPf], qha[kPf], qe[kPf], qeh[kPf];
        float qf1[kPf], qf1h[kPf], qf2[kPf], qf2h[kPf], qm[kPf];
        auto fill = [&](int i, int y) {
            qb[i] = bload(rb, xo, row_off(y + 1));
            qq[i] = bload(rb, qo, row_off(y + 1));
            qk[i] = bload(rk, xo, row_off(y + 1));
            qkq[i] = bload(rk, qo, row_off(y + 1));
            qa[i] = bload(ra, xo, row_off(y));
            qha[i] = bload(ra, ho, row_off(y));
            qe[i] = 0.0f;
            qeh[i] = 0.0f;
            if (inject) {
                qe[i] = bload(re, xo, row_off(y));
                qeh[i] = bload(re, ho, row_off(y));
            }
            qf1[i] = qf1h[i] = qf2[i] = qf2h[i] = 0.0f;
            if constexpr (IMG) {
                qf1[i] = bload(rf1, xo, row_off(y));
                qf1h[i] = bload(rf1, ho, row_off(y));
                qf2[i] = bload(rf2, xo, row_off(y - 1));
                qf2h[i] = bload(rf2, ho, row_off(y - 1));
            }
            qm[i] = 0.0f;
            if constexpr (EDGE) qm[i] = bload(rm, mo, mrow(y));
        };
#pragma unroll
        for (int i = 0; i < kPf; ++i) {
            __builtin_amdgcn_sched_barrier(0);
            fill(i, y_start + i);
        }
        __builtin_amdgcn_sched_barrier(0);
        float cm2 = 0.0f, cm1 = 0.0f, hcm1 = 0.0f, pcm2 = 0.0f, pcm1 = 0.0f, hpcm1 = 0.0f;
        float dm1 = 0.0f, sm1 = 0.0f, t1_prev = 0.0f;
        int em1 = -1;
        auto row = [&](int y, int i) {
            const bool row_in = y < H;
            const bool nx_in = y + 1 < H;
            float bp = (xin & nx_in) ? qb[i] : 0.0f;
            float qp = (xq_in & nx_in) ? qq[i] : 0.0f;
            float kp = (xin & nx_in) ? qk[i] : 0.0f;
            float kqp = (xq_in & nx_in) ? qkq[i] : 0.0f;
            float a0 = (xin & row_in) ? qa[i] : 0.0f;
            float ha0 = (xh_in & row_in) ? qha[i] : 0.0f;
            int e0 = (xin & row_in) ? __float_as_int(qe[i]) : -1;
            int eh0 = (xh_in & row_in) ? __float_as_int(qeh[i]) : -1;
            float f10 = (xin & row_in) ? qf1[i] : 0.0f;
            float f1h0 = (xh_in & row_in) ? qf1h[i] : 0.0f;
            const bool f2_in = (y - 1 >= 0) & (y - 1 < H);
            float f2m1 = (xin & f2_in) ? qf2[i] : 0.0f;
            float f2hm1 = (xh_in & f2_in) ? qf2h[i] : 0.0f;
            float s0 = qm[i];
            asm volatile("" : "+v"(bp), "+v"(qp), "+v"(kp), "+v"(kqp), "+v"(a0), "+v"(ha0),
                         "+v"(e0), "+v"(eh0), "+v"(f10), "+v"(f1h0), "+v"(f2m1), "+v"(f2hm1),
                         "+v"(s0));
            __builtin_amdgcn_sched_barrier(0);
            // rows past y_end are not needed: reload row y_end instead (a cache
            // hit; a conditional refill would make every wait a full one)
            fill(i, min(y + kPf, y_end));
            // ---- first step (step n), row y: block columns, then halo columns
            const float P0 = k0 * b0;
            const float ph0 = kq0 * q0;
            float c0 = 0.0f, hc = 0.0f;
            if (row_in) {
                const float Pm = km * bm, Pp = kp * bp;
                const float pw = from_west(P0, ph0), pe = from_east(P0, ph0);
                c0 = __fmaf_rn(dt2, lap5(pw, pe, Pm, Pp, P0, p.idx2, p.idy2, p.cen),
                               leap(b0, a0));
                if (inject && e0 >= 0) c0 += injected(res1, e0);
                if constexpr (EDGE) {
                    if (y >= 1 && y <= H - 2) c0 = mur_sides(c0, b0, x, W, s0);
                    if (y == H - 1) c0 = mur_row(mur, W, H, x, kt_x, b0, bm, cm1, 2);
                }
                const float phe = from_east(ph0, 0.0f), phw = from_west(ph0, 0.0f);
                const float hw = hl ? phe : P0;
                const float he = hl ? P0 : phw;
                hc = __fmaf_rn(dt2, lap5(hw, he, kqm1 * qm1, kqp * qp, ph0, p.idx2, p.idy2,
                                         p.cen),
                               leap(q0, ha0));
                if (inject && eh0 >= 0) hc += injected(res1, eh0);
                if constexpr (EDGE) {
                    if (h_right_edge) {
                        if (y >= 1 && y <= H - 2) {
                            hc = b0 + s0 * (c0 - q0);
                        } else if (y == H - 1) {
                            const float from_x = b0 + kx3 * (c0 - q0);
                            const float from_y = qm1 + ky3 * (hcm1 - q0);
                            hc = 0.5f * (from_x + from_y);
                        }
                    } else if (y == H - 1) {
                        hc = qm1 + kt_h * (hcm1 - q0);
                    }
                    if (y == 1) {
                        cm1 = mur_row(mur, W, H, x, kb_x, bm, b0, c0, 0);
                        if (h_right_edge) {
                            const float from_x = bm + kx1 * (cm1 - qm1);
                            const float from_y = q0 + ky1 * (hc - qm1);
                            hcm1 = 0.5f * (from_x + from_y);
                        } else {
                            hcm1 = q0 + kb_h * (hc - qm1);
                        }
                        pcm1 = km * cm1;
                        hpcm1 = kqm1 * hcm1;
                        if (owned & (r0 == 0)) outc[xo / 4] = cm1;
                    }
                }
                if (owned & (y >= r0) & (y < r1) & (y > 0)) outc[(size_t)y * W + x] = c0;
            }
            const float pc0 = k0 * c0;
            const float hpc0 = kq0 * hc;
            // ---- second step (step n-1), row r = y - 1
            const int r = y - 1;
            float d = 0.0f;
            if (r >= r0 && r < r1) {
                const float pw = from_west(pcm1, hpcm1), pe = from_east(pcm1, hpcm1);
                d = __fmaf_rn(dt2, lap5(pw, pe, pcm2, pc0, pcm1, p.idx2, p.idy2, p.cen),
                              leap(cm1, bm));
                if (inject && em1 >= 0) d += injected(res2, em1);
                if constexpr (EDGE) {
                    if (r >= 1 && r <= H - 2) d = mur_sides(d, cm1, x, W, sm1);
                    if (r == H - 1) d = mur_row(mur, W, H, x, kt_x, cm1, cm2, dm1, 2);
                    if (r == 1) {
                        const float d0 = mur_row(mur, W, H, x, kb_x, cm2, cm1, d, 0);
                        if (owned) outd[xo / 4] = d0;
                        dm1 = d0;  // row 0, final, for its imaging below
                    }
                }
                if (owned & (r > 0)) outd[(size_t)r * W + x] = d;
            }
            // ---- imaging: C of row y-1 (final now) and D of row y-2
            if constexpr (IMG) {
                float t1 = 0.0f, t2 = 0.0f;
                if (use1) {
                    const float fw = from_west(f1m1, f1hm1), fe = from_east(f1m1, f1hm1);
                    t1 = cm1 * lap5(fw, fe, f1m2, f10, f1m1, p.idx2, p.idy2, p.cen);
                }
                if (use2) {
                    const float fw = from_west(f2m2, f2hm2), fe = from_east(f2m2, f2hm2);
                    t2 = dm1 * lap5(fw, fe, f2m3, f2m1, f2m2, p.idx2, p.idy2, p.cen);
                }
                const int ra2 = y - 2;
                if (ra2 >= r0 && ra2 < r1) {
                    float *cell = acc + (ra2 - r0) * kLanes + lane;
                    *cell += t1_prev + t2;
                }
                t1_prev = t1;
                f1m2 = f1m1;
                f1m1 = f10;
                f1hm1 = f1h0;
                f2m3 = f2m2;
                f2m2 = f2m1;
                f2hm2 = f2hm1;
            }
            cm2 = cm1;
            cm1 = c0;
            hcm1 = hc;
            pcm2 = pcm1;
            pcm1 = pc0;
            hpcm1 = hpc0;
            bm = b0;
            b0 = bp;
            qm1 = q0;
            q0 = qp;
            km = k0;
            k0 = kp;
            kqm1 = kq0;
            kq0 = kqp;
            dm1 = d;
            sm1 = s0;
            em1 = e0;
        };
        for (int y = y_start; y <= y_end; y += kPf) {
#pragma unroll
            for (int i = 0; i < kPf; ++i) row(y + i, i);
        }
    }
}

extern __shared__ float adj2_acc[];

// Workgroup b of a two-step adjoint launch of one problem.
template <unsigned F>
__device__ __forceinline__ void adj2_block(const StepParams &p, int b)
{
    constexpr bool IMG = (F & kImage) != 0;
    const int wblk = (b % kXcds) * p.chunk + b / kXcds;
    if (wblk >= p.n_work) return;  // uniform over the workgroup
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kLanes));
    const int grp = wblk % p.n_groups;
    const int t = wblk / p.n_groups;
    const int cb = t % p.n_cb;
    const int seg = t / p.n_cb;
    const int xs = min(cb * kLanes, max(p.nx - kLanes, 0));
    int r0, r1;
    seg_rows(p, seg, r0, r1);
    const bool edge = (xs <= 1) || (xs + kLanes >= p.nx - 1) || (r0 == 0) || (r1 >= p.ny - 1);
    float *acc = adj2_acc + (size_t)wave * p.acc_rows * kLanes;
    if (edge)
        adj2_wave<true, IMG>(p, lane, wave, grp, cb, seg, acc);
    else
        adj2_wave<false, IMG>(p, lane, wave, grp, cb, seg, acc);
    if constexpr (IMG) {
        if (p.img == nullptr && p.img_b == nullptr) return;  // uniform
        __syncthreads();
        // partial[grp] += scale * (wave 0 + wave 1 + wave 2 + wave 3), per cell
        const int W = p.nx;
        const int own0 = cb * kLanes;
        const int x = xs + lane;
        const bool owned = (x < W) & (x >= own0);
        gf *gp = gstore_ptr(p.g_part) + (size_t)grp * ((size_t)p.ny * W);
        const int n_rows = r1 - r0;
        for (int i = wave; i < n_rows; i += kWavesPerBlock) {
            float sum = adj2_acc[i * kLanes + lane];
#pragma unroll
            for (int w = 1; w < kWavesPerBlock; ++w)
                sum += adj2_acc[((size_t)w * p.acc_rows + i) * kLanes + lane];
            if (owned) {
                gf *cell = gp + (size_t)(r0 + i) * W + x;
                *cell = __fmaf_rn(p.img_scale, sum, *cell);
            }
        }
    }
}

template <unsigned F>
__global__ void __launch_bounds__(256) adj2_kernel(StepParams p)
{
    adj2_block<F>(p, (int)blockIdx.x);
}

// Batched two-step adjoint launch (see fwd2_batch_kernel); the LDS of the
// launch is sized for the largest imaging accumulator of its entries.
template <unsigned F>
__global__ void __launch_bounds__(256) adj2_batch_kernel(const StepParams *table, const int *first,
                                                         int n)
{
    const int b = (int)blockIdx.x;
    const int e = __builtin_amdgcn_readfirstlane(batch_entry(first, n, b));
    adj2_block<F>(batch_params(table, e), b - first[e]);
}

// ---------------------------------------------------------------------------
// Several adjoint steps per launch, full-width workgroups
// ---------------------------------------------------------------------------
//
// Steps n, n-1, ..., n-KL+1 of the reverse run in one launch, with the
// full-width workgroups and the level pipeline of the forward kernel above:
// level j (1..KL) computes lam' = 2 lam - lam_old + dt^2 Lap(c^2 lam) +
// residual, then the forward Mur update, at row y - j + 1 of iteration y.
// Each level keeps lam and P = c^2 lam of its three newest rows; the edge
// lanes exchange P (and the imaging fields) through LDS. The arithmetic per
// cell and step is the single-step kernel's, so the states are the same bits.
//
// Imaging: level j (step n - j + 1) is imaged against the forward field of
// its step, lv_in[j - 1] (zero outside the grid), one row late so that row 0
// is final: g += scale sum_s lam_j Lap(F_j). A workgroup takes one row
// segment of one shot group and marches the group's shots one after the
// other; a cell's terms of the KL levels are summed in registers and then
// added into an LDS image of the segment (each lane owns its column, so the
// order is fixed); at the end the workgroup adds the segment into the group's
// partial image (at most max_wide_groups), which are summed in group order
// after the run: the gradient is deterministic.

constexpr int kPfA = 2;  // rows loaded ahead in the adjoint march

// col[i W] += scale * acc[i acc_stride] for rows i < n_rows of a lane's
// column of a partial image, one fused multiply-add per cell. The rows'
// loads go out kImgBatch at a time before their stores: a row's load cannot
// move above the previous row's store (the compiler cannot prove the rows
// apart), so one row at a time would wait a memory latency per row.
constexpr int kImgBatch = 16;
__device__ __forceinline__ void add_image_column(gf *col, int n_rows, int W, float scale,
                                                 const float *acc, int acc_stride)
{
    int i = 0;
    for (; i + kImgBatch <= n_rows; i += kImgBatch) {
        float v[kImgBatch];
#pragma unroll
        for (int k = 0; k < kImgBatch; ++k) v[k] = col[(size_t)(i + k) * W];
#pragma unroll
        for (int k = 0; k < kImgBatch; ++k)
            col[(size_t)(i + k) * W] = __fmaf_rn(scale, acc[(i + k) * acc_stride], v[k]);
    }
    for (; i < n_rows; ++i)
        col[(size_t)i * W] = __fmaf_rn(scale, acc[i * acc_stride], col[(size_t)i * W]);
}

// (P of level j - 1, imaging field of level j) for j = 1..KL
struct alignas(16) AdjEdge {
    float2 v[kMaxLevels];
};
typedef AdjEdge AdjEdgeRow[kMaxWideWaves + 2][2];

extern __shared__ float adjm_acc[];

// IMG: 0 no imaging, 1 imaging against fields (levels without one skip it),
// 2 against their Laplacians, 3 against fields with every level imaged.
template <int KL, bool XE, bool YE, int IMG>
__device__ __forceinline__ void adjm_wave(const StepParams &p, int lane, int wave, int grp,
                                          int seg, AdjEdgeRow *ex)
{
    // imaging against the fields themselves (3: every level has one)
    constexpr bool FIELD = IMG == 1 || IMG == 3;
    const int W = p.nx, H = p.ny;
    const int x = wave * kLanes - p.wide_off + lane;
    // only the first and last wavefronts of a row (XE) reach past the grid
    const bool owned = !XE || ((x >= 0) & (x < W));
    const int xo = min(max(x, 0), W - 1) * 4;
    const int so = owned ? x * 4 : kDropOffset;
    int r0, r1;
    seg_rows(p, seg, r0, r1);
    const int y_begin = max(r0 - (KL - 1), 0);
    // level KL reaches row r1 - 1; imaging one row late needs one more
    const int y_end = r1 + KL - 2 + (IMG != 0 ? 1 : 0);
    const size_t plane = (size_t)H * W;
    const unsigned plane_bytes = (unsigned)(plane * sizeof(float));
    const int row_bytes = W * 4;
    // segments away from the top and bottom (!YE) load only rows inside the grid
    auto row_off = [&](int yy) -> int {
        if constexpr (YE) return min(max(yy, 0), H - 1) * row_bytes;
        else return yy * row_bytes;
    };
    const __amdgpu_buffer_rsrc_t rk = make_rsrc(p.c2, plane_bytes);
    // the scatter-table map, Mur side coefficients and imaging fields are read
    // with global loads off a scalar base (2 SGPRs each instead of a 4-SGPR
    // buffer descriptor: the adjoint is short of SGPRs)
    const float *emap = (const float *)p.cell_map;
    const float *mur = p.mur;
    const int mo = (x == 0) ? 0 : H * 4;
    auto mrow = [&](int yy) -> int {
        if constexpr (YE) return min(max(yy, 0), H - 1) * 4;
        else return yy * 4;
    };
    const float dt2 = p.dt2;
    float kb_x = 0.0f, kt_x = 0.0f;
    if constexpr (YE) {
        kb_x = mur[2 * H + xo / 4];
        kt_x = mur[2 * H + W + xo / 4];
    }
    const int w_slot = lane == 0 ? wave : wave + 1;
    const int w_side = lane == 0 ? 1 : 0;
    const int e_slot = lane == kLanes - 1 ? wave + 2 : wave + 1;
    const int e_side = lane == kLanes - 1 ? 0 : 1;
    const bool edge_lane = (lane == 0) | (lane == kLanes - 1);
    const int my_side = lane == kLanes - 1 ? 1 : 0;
    const int s_begin = grp * p.shots_per_group;
    const int s_end = min(p.n_shots, s_begin + p.shots_per_group);
    float *acc = adjm_acc;
    for (int s = s_begin; s < s_end; ++s) {
        const size_t soff = (size_t)s * plane;
        const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.f_old + soff, plane_bytes);
        const __amdgpu_buffer_rsrc_t rb = make_rsrc(p.f_cur + soff, plane_bytes);
        __amdgpu_buffer_rsrc_t ro[KL];
        const float *rf[KL];
#pragma unroll
        for (int j = 0; j < KL; ++j) {
            // integer arithmetic: the pointers may be NULL (not stored / not imaged)
            ro[j] = make_rsrc((const float *)((uintptr_t)p.lv_out[j] + soff * sizeof(float)),
                              plane_bytes);
            // a level without imaging field reads (and ignores) the state instead
            rf[j] = (p.lv_in[j] != nullptr ? p.lv_in[j] : p.f_cur) + soff;
        }
        // residual rows of the levels: formed only when a lane injects (rare),
        // to keep scalar registers free
        const size_t res_off = (size_t)s * p.n_rec;
        // lam of each level's newest three rows (level 0 = the input state,
        // newest row y + 1 at iteration y); P = c^2 lam is formed where it is
        // needed (the same product each time, so the same bits), which keeps
        // the register budget down
        float lv[KL + 1][3];
        // c^2 of rows y + 1, y, ..., y - KL; scatter entry and Mur side
        // coefficient of rows y, ..., y - KL + 1; imaging field of level j at
        // rows y - j + 1, y - j, y - j - 1
        float kh[KL + 2], mh[KL], fh[KL + 1][3];
        int eh[KL];
        // imaging terms of rows y - 1, ..., y - KL summed over the levels so
        // far (each cell: ((t_1 + t_2) + ...) + t_KL, then into the LDS image)
        float racc[KL];
#pragma unroll
        for (int j = 0; j < KL; ++j) racc[j] = 0.0f;
#pragma unroll
        for (int j = 0; j <= KL; ++j)
            lv[j][0] = lv[j][1] = lv[j][2] = fh[j][0] = fh[j][1] = fh[j][2] = 0.0f;
#pragma unroll
        for (int j = 0; j < KL + 2; ++j) kh[j] = 0.0f;
#pragma unroll
        for (int j = 0; j < KL; ++j) {
            mh[j] = 0.0f;
            eh[j] = -1;
        }
        lv[0][0] = bload(rb, xo, row_off(y_begin));
        lv[0][1] = bload(rb, xo, row_off(y_begin - 1));
        kh[0] = bload(rk, xo, row_off(y_begin));
        kh[1] = bload(rk, xo, row_off(y_begin - 1));
        float qb[kPfA], qa[kPfA], qk[kPfA], qe[kPfA], qm[kPfA], qf[kPfA][KL];
        auto fill = [&](int i, int y) {
            qb[i] = bload(rb, xo, row_off(y + 1));
            qa[i] = bload(ra, xo, row_off(y));
            qk[i] = bload(rk, xo, row_off(y + 1));
            // every iteration issues the same loads, so the compiler's waits
            // count exactly (a conditional load makes them wait for loads
            // issued after it)
            qe[i] = gload_off(emap, xo + row_off(y));
            qm[i] = 0.0f;
            if constexpr (XE) qm[i] = gload_off(mur, mo + mrow(y));
#pragma unroll
            for (int j = 0; j < KL; ++j) {
                // level j + 1: newest field row (y - j), or the Laplacian of the
                // row it images (y - j - 1)
                qf[i][j] = 0.0f;
                if constexpr (FIELD) qf[i][j] = gload_off(rf[j], xo + row_off(y - j));
                if constexpr (IMG == 2) qf[i][j] = gload_off(rf[j], xo + row_off(y - j - 1));
            }
        };
#pragma unroll
        for (int i = 0; i < kPfA; ++i) {
            __builtin_amdgcn_sched_barrier(0);
            fill(i, y_begin + i);
        }
        __builtin_amdgcn_sched_barrier(0);
        if (edge_lane) {
            AdjEdge e;
            e.v[0] = make_float2(kh[0] * lv[0][0], 0.0f);
#pragma unroll
            for (int j = 1; j < kMaxLevels; ++j) e.v[j] = make_float2(0.0f, 0.0f);
            ex[y_begin & 1][wave + 1][my_side] = e;
        }
        __syncthreads();
        auto iteration = [&](int y, int i) {
#pragma unroll
            for (int j = 0; j <= KL; ++j) {
                lv[j][2] = lv[j][1];
                lv[j][1] = lv[j][0];
                fh[j][2] = fh[j][1];
                fh[j][1] = fh[j][0];
            }
#pragma unroll
            for (int j = KL + 1; j >= 1; --j) kh[j] = kh[j - 1];
#pragma unroll
            for (int j = KL - 1; j >= 1; --j) {
                mh[j] = mh[j - 1];
                eh[j] = eh[j - 1];
            }
            const float b1 = take(qb[i]), a0 = take(qa[i]), k1 = take(qk[i]), e0 = take(qe[i]);
            const float m0 = XE ? take(qm[i]) : 0.0f;
            float f[KL];
#pragma unroll
            for (int j = 0; j < KL; ++j) f[j] = FIELD || IMG == 2 ? take(qf[i][j]) : 0.0f;
            __builtin_amdgcn_sched_barrier(0);
            fill(i, min(y + kPfA, y_end));
            lv[0][0] = b1;
            kh[0] = k1;
            eh[0] = owned ? __float_as_int(e0) : -1;
            mh[0] = m0;
#pragma unroll
            for (int j = 1; j <= KL; ++j) {
                if constexpr (FIELD) {
                    const int fr = y - j + 1;  // row of level j's newest imaging field
                    // segments away from the top and bottom (!YE) image inside the grid
                    fh[j][0] = (owned & (!YE || ((fr >= 0) & (fr < H)))) ? f[j - 1] : 0.0f;
                } else {
                    fh[j][0] = f[j - 1];
                }
            }
#pragma unroll
            for (int j = 1; j <= KL; ++j) {
                const int rho = y - j + 1;
                const AdjEdge *hw = &ex[y & 1][w_slot][w_side];
                const AdjEdge *he = &ex[y & 1][e_slot][e_side];
                const float u = lv[j - 1][1];
                const float old = j == 1 ? a0 : lv[j - 2][2];
                // P of the input level at rows rho - 1, rho, rho + 1
                const float Ps = kh[j + 1] * lv[j - 1][2];
                const float P0 = kh[j] * u;
                const float Pn = kh[j - 1] * lv[j - 1][0];
                const float pw = from_west(P0, hw->v[j - 1].x), pe = from_east(P0, he->v[j - 1].x);
                float v = __fmaf_rn(dt2, lap5(pw, pe, Ps, Pn, P0, p.idx2, p.idy2, p.cen),
                                    leap(u, old));
                const int e = eh[j - 1];
                v = inject_lanes(p, p.lv_resid[j - 1], res_off, e, lane, v);
                if constexpr (XE) {
                    if (!YE || (rho >= 1 && rho <= H - 2)) v = mur_sides(v, u, x, W, mh[j - 1]);
                }
                if constexpr (YE) {
                    if (rho == H - 1) v = mur_row(mur, W, H, x, kt_x, u, lv[j - 1][2], lv[j][1], 2);
                    if (rho == 1) {
                        const float v0 = mur_row(mur, W, H, x, kb_x, lv[j - 1][2], u, v, 0);
                        lv[j][1] = v0;
                        if ((j >= KL - 1) & (r0 == 0)) bstore(ro[j - 1], v0, so, 0);
                    }
                }
                lv[j][0] = v;
                if (j >= KL - 1) {
                    // stores of rows outside the segment are dropped, not
                    // branched around (the same operations every iteration)
                    const bool st = (rho >= r0) & (rho < r1) & (rho > 0);
                    bstore(ro[j - 1], v, st ? so : kDropOffset, st ? rho * row_bytes : 0);
                }
                if constexpr (IMG != 0) {
                    // row rho - 1 = y - j: its term goes to racc[j - 1]
                    if (IMG == 3 || p.lv_in[j - 1] != nullptr) {
                        float t;
                        if constexpr (FIELD) {
                            const float F0 = fh[j][1];
                            const float fw = from_west(F0, hw->v[j - 1].y),
                                        fe = from_east(F0, he->v[j - 1].y);
                            t = lv[j][1] * lap5(fw, fe, fh[j][2], fh[j][0], F0, p.idx2, p.idy2,
                                                p.cen);
                        } else {
                            t = lv[j][1] * fh[j][0];  // the Laplacian of row rho - 1
                        }
                        racc[j - 1] += t;
                    }
                }
            }
            if constexpr (IMG != 0) {
                // row y - KL has the terms of every level: add it to the image
                const int Rf = y - KL;
                if ((Rf >= r0) & (Rf < r1) & owned) acc[(Rf - r0) * W + x] += racc[KL - 1];
#pragma unroll
                for (int k = KL - 1; k >= 1; --k) racc[k] = racc[k - 1];
                racc[0] = 0.0f;
            }
            if (edge_lane) {
                AdjEdge e;
#pragma unroll
                for (int j = 0; j < kMaxLevels; ++j)
                    e.v[j] = j < KL ? make_float2(kh[j] * lv[j][0], FIELD ? fh[j + 1][0] : 0.0f)
                                    : make_float2(0.0f, 0.0f);
                ex[(y + 1) & 1][wave + 1][my_side] = e;
            }
            __syncthreads();
        };
        for (int y = y_begin; y <= y_end; y += kPfA) {
#pragma unroll
            for (int i = 0; i < kPfA; ++i) iteration(y + i, i);
        }
    }
}

// Workgroup b (row segment, shot group) of a launch of KL adjoint steps.
template <int KL, int IMG>
__device__ __forceinline__ void adjm_block(const StepParams &p, int b, AdjEdgeRow *ex)
{
    const int wblk = (b % kXcds) * p.chunk + b / kXcds;
    if (wblk >= p.n_work) return;  // uniform over the workgroup
    const int t = (int)threadIdx.x;
    if (t < 2 * 2 * kMaxLevels) {  // zero pads at both ends, both parities
        const int par = t / (2 * kMaxLevels), side = (t / kMaxLevels) & 1, j = t % kMaxLevels;
        ex[par][0][side].v[j] = make_float2(0.0f, 0.0f);
        ex[par][p.wide_waves + 1][side].v[j] = make_float2(0.0f, 0.0f);
    }
    const int lane = t & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane(t / kLanes);
    const int grp = wblk % p.n_groups;
    const int seg = wblk / p.n_groups;
    int r0, r1;
    seg_rows(p, seg, r0, r1);
    const int W = p.nx;
    const int n_acc = (r1 - r0) * W;
    if constexpr (IMG != 0) {
        for (int i = t; i < n_acc; i += (int)blockDim.x) adjm_acc[i] = 0.0f;
    }
    // (the first barrier of the march orders these writes before any use)
    const bool xe = (wave == 0) || (wave == p.wide_waves - 1);
    // rows the march loads: y_begin - KL .. r1 + KL (imaging fields KL - 1 or KL
    // rows behind, inputs one ahead); segments whose loads stay inside the grid
    // take the variant without clamps or Mur row updates
    const bool ye = (r0 < 2 * KL) || (r1 + KL >= p.ny);
    if (xe) {
        if (ye) adjm_wave<KL, true, true, IMG>(p, lane, wave, grp, seg, ex);
        else adjm_wave<KL, true, false, IMG>(p, lane, wave, grp, seg, ex);
    } else {
        if (ye) adjm_wave<KL, false, true, IMG>(p, lane, wave, grp, seg, ex);
        else adjm_wave<KL, false, false, IMG>(p, lane, wave, grp, seg, ex);
    }
    if constexpr (IMG != 0) {
        __syncthreads();
        // partial[grp] += scale * segment image, per cell
        const int x = wave * kLanes - p.wide_off + lane;
        if ((x >= 0) & (x < W)) {
            gf *gp = gstore_ptr(p.g_part) + (size_t)grp * ((size_t)p.ny * W);
            add_image_column(gp + (size_t)r0 * W + x, r1 - r0, W, p.img_scale, adjm_acc + x, W);
        }
    }
}

template <int KL, int IMG>
__global__ void __launch_bounds__(1024) adjm_kernel(StepParams p)
{
    __shared__ AdjEdgeRow ex[2];
    adjm_block<KL, IMG>(p, (int)blockIdx.x, ex);
}

// Batched full-width adjoint launch (entries of one row width; the LDS is
// sized for the largest segment image of the entries).
template <int KL, int IMG>
__global__ void __launch_bounds__(1024) adjm_batch_kernel(const StepParams *table,
                                                          const int *first, int n)
{
    __shared__ AdjEdgeRow ex[2];
    const int b = (int)blockIdx.x;
    const int e = __builtin_amdgcn_readfirstlane(batch_entry(first, n, b));
    adjm_block<KL, IMG>(batch_params(table, e), b - first[e], ex);
}

// ---------------------------------------------------------------------------
// Adjoint launches that rebuild the forward field backwards in time
// ---------------------------------------------------------------------------
//
// The leapfrog is reversible in the interior: U(m-2) = 2 U(m-1) - U(m) +
// dt^2 c^2 Lap(U(m-1)) + s(m), s(m) the transmitter's value of step m. The Mur
// update of the boundary cells is not (solved for the older value it divides
// by |k| < 1 every step), so the forward run keeps the boundary cells of every
// step in a ring and the reconstruction takes them from there. Adjoint step m
// images lam' against Lap(U(m-1)), the Laplacian that the reconstruction of
// U(m-2) takes anyway: a launch of KL adjoint steps reads and writes a pair of
// forward fields instead of reading KL stored ones, and no forward
// recomputation is needed. The reconstruction rounds differently from the
// forward run (about 1e-6 relative after thousands of steps); the adjoint
// states are those of adjm_kernel bit for bit.
//
// Level j (1..KL) of a launch whose first step is n runs adjoint step
// m = n - j + 1 exactly as adjm_kernel does and reconstructs uv[j] = U(n-j-1)
// in the same row pipeline: at iteration y it takes row y - j + 1 from the
// three newest rows of uv[j - 1] (uv[0] = U(n-1), marched as rows) and the row
// of uv[j - 2] (U(n) for level 1, read per row). The imaging term of row
// y - j (one row late, as in adjm_kernel) uses the Laplacian kept from the
// previous iteration. Lanes and rows outside the grid hold zeros, so the
// imaging Laplacian sees the zero padding the stored fields had. Boundary
// cells come from the ring: the two columns staged in LDS per shot (rows of
// the segment plus the pipeline halo, every level), rows 0 and ny - 1 read
// when the level reaches them.

// (first adjoint step of the launch, level j) -> index of the reconstructed
// field U(n - j - 1)
__device__ __forceinline__ int recon_field(int n, int j) { return n - j - 1; }

// Boundary columns of shot s's rebuilt levels for a segment whose march
// starts at row yb + KL - 1, staged in LDS: bcol[(j * 2 + side) * nb + row -
// yb] for level j + 1, side 0 = column 0, 1 = column W - 1. Every thread of the
// workgroup takes part; the barrier before the march orders the writes.
template <int KL>
__device__ __forceinline__ void adjr_stage_columns(const StepParams &p, int s, int yb,
                                                   float *bcol)
{
    const int H = p.ny, W = p.nx, nb = p.bcol_rows;
    const size_t bnd_n = bnd_floats(H, W);
    for (int i = (int)threadIdx.x; i < KL * 2 * nb; i += (int)blockDim.x) {
        const int r = i % nb, side = (i / nb) & 1, j = i / (2 * nb);
        const int row = yb + r;
        const int q = recon_field(p.src_index, j + 1);
        const int k = q - p.bnd_first;
        float v = 0.0f;
        if ((row >= 1) & (row <= H - 2) & (q >= 0) & (k >= 0) & (k < p.bnd_slots))
            v = p.bnd[((size_t)k * p.n_shots + s) * bnd_n + 2 * W + (size_t)side * (H - 2) +
                      row - 1];
        bcol[i] = v;
    }
}

template <int KL, bool XE, bool YE, bool EVERY>
__device__ __forceinline__ void adjr_wave(const StepParams &p, int lane, int wave, int grp,
                                          int seg, AdjEdgeRow *ex, float *bcol, float *brow)
{
    // staged rows: [0][level][x] row 0, [brow_second][level][x] row H - 1 (the
    // same slot unless a segment's march can reach both: H <= rows + 3 KL)
    const int brow_second = p.ny <= p.acc_rows + 3 * KL ? 1 : 0;
    const int W = p.nx, H = p.ny;
    const int x = wave * kLanes - p.wide_off + lane;
    const bool owned = !XE || ((x >= 0) & (x < W));
    const int xo = min(max(x, 0), W - 1) * 4;
    const int so = owned ? x * 4 : kDropOffset;
    int r0, r1;
    seg_rows(p, seg, r0, r1);
    const int y_begin = max(r0 - (KL - 1), 0);
    // level KL reaches row r1 - 1; imaging one row late needs one more
    const int y_end = r1 + KL - 1;
    const size_t plane = (size_t)H * W;
    const unsigned plane_bytes = (unsigned)(plane * sizeof(float));
    const int row_bytes = W * 4;
    auto row_off = [&](int yy) -> int {
        if constexpr (YE) return min(max(yy, 0), H - 1) * row_bytes;
        else return yy * row_bytes;
    };
    // forward-field rows outside the grid are zero
    auto in_rows = [&](int yy) -> bool {
        if constexpr (YE) return (yy >= 0) & (yy < H);
        else return true;
    };
    const __amdgpu_buffer_rsrc_t rk = make_rsrc(p.c2, plane_bytes);
    const float *emap = (const float *)p.cell_map;
    const float *mur = p.mur;
    const int mo = (x == 0) ? 0 : H * 4;
    auto mrow = [&](int yy) -> int {
        if constexpr (YE) return min(max(yy, 0), H - 1) * 4;
        else return yy * 4;
    };
    const float dt2 = p.dt2;
    float kb_x = 0.0f, kt_x = 0.0f;
    if constexpr (YE) {
        kb_x = mur[2 * H + xo / 4];
        kt_x = mur[2 * H + W + xo / 4];
    }
    const int w_slot = lane == 0 ? wave : wave + 1;
    const int w_side = lane == 0 ? 1 : 0;
    const int e_slot = lane == kLanes - 1 ? wave + 2 : wave + 1;
    const int e_side = lane == kLanes - 1 ? 0 : 1;
    const bool edge_lane = (lane == 0) | (lane == kLanes - 1);
    const int my_side = lane == kLanes - 1 ? 1 : 0;
    const int n_top = p.src_index;
    const size_t bnd_n = bnd_floats(H, W);
    // staged boundary columns: bcol[(j * 2 + side) * nb + row - yb] for level
    // j + 1, side 0 = column 0, 1 = column W - 1
    const int yb = y_begin - KL + 1;
    const int nb = p.bcol_rows;
    const int col_side = (x == 0) ? 0 : 1;
    const bool side_lane = (x == 0) | (x == W - 1);
    const int s_begin = grp * p.shots_per_group;
    const int s_end = min(p.n_shots, s_begin + p.shots_per_group);
    float *acc = adjm_acc;
    for (int s = s_begin; s < s_end; ++s) {
        const size_t soff = (size_t)s * plane;
        adjr_stage_columns<KL>(p, s, yb, bcol);
        // rows 0 and H - 1 of the levels, for segments that reach them (staged
        // too: a conditional vector load in the march would make the compiler
        // wait for every outstanding load after it)
        if constexpr (YE) {
            for (int which = 0; which < 2; ++which) {
                const bool need = which == 0 ? (yb <= 0) : (y_end + 1 >= H - 1);
                if (!need) continue;  // uniform
                float *dst = brow + (size_t)(which == 0 ? 0 : brow_second) * KL * W;
                for (int i = (int)threadIdx.x; i < KL * W; i += (int)blockDim.x) {
                    const int xx = i % W, j = i / W;
                    const int q = recon_field(n_top, j + 1);
                    const int k = q - p.bnd_first;
                    float v = 0.0f;
                    if ((q >= 0) & (k >= 0) & (k < p.bnd_slots))
                        v = p.bnd[((size_t)k * p.n_shots + s) * bnd_n + (size_t)which * W + xx];
                    dst[i] = v;
                }
            }
        }
        const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.f_old + soff, plane_bytes);
        const __amdgpu_buffer_rsrc_t rb = make_rsrc(p.f_cur + soff, plane_bytes);
        const __amdgpu_buffer_rsrc_t rua = make_rsrc(p.r_next + soff, plane_bytes);
        const __amdgpu_buffer_rsrc_t rub = make_rsrc(p.r_cur + soff, plane_bytes);
        __amdgpu_buffer_rsrc_t ro[KL], rou[KL];
#pragma unroll
        for (int j = 0; j < KL; ++j) {
            // integer arithmetic: the pointers may be NULL (not stored)
            ro[j] = make_rsrc((const float *)((uintptr_t)p.lv_out[j] + soff * sizeof(float)),
                              plane_bytes);
            rou[j] = make_rsrc((const float *)((uintptr_t)p.lv_rout[j] + soff * sizeof(float)),
                               plane_bytes);
        }
        const size_t res_off = (size_t)s * p.n_rec;
        const size_t res_step = (size_t)p.n_shots * p.n_rec;
        const int sr = p.src_rows[s], sc = p.src_cols[s];
        const int src_extra = transmitter_rows(p);
        // lam and U of each level's newest three rows (level 0: the inputs,
        // newest row y + 1 at iteration y); P = c^2 lam formed where needed
        float lv[KL + 1][3], uv[KL + 1][3];
        // c^2 of rows y + 1, ..., y - KL; scatter entry and Mur side
        // coefficient of rows y, ..., y - KL + 1
        float kh[KL + 2], mh[KL];
        int eh[KL];
        // imaging terms of rows y - 1 .. y - KL summed over the levels so far,
        // and each level's Laplacian of the previous iteration's row
        float racc[KL], lsave[KL];
#pragma unroll
        for (int j = 0; j < KL; ++j) {
            racc[j] = lsave[j] = 0.0f;
            mh[j] = 0.0f;
            eh[j] = -1;
        }
#pragma unroll
        for (int j = 0; j <= KL; ++j)
            lv[j][0] = lv[j][1] = lv[j][2] = uv[j][0] = uv[j][1] = uv[j][2] = 0.0f;
#pragma unroll
        for (int j = 0; j < KL + 2; ++j) kh[j] = 0.0f;
        lv[0][0] = bload(rb, xo, row_off(y_begin));
        lv[0][1] = bload(rb, xo, row_off(y_begin - 1));
        {
            const float u0 = bload(rub, xo, row_off(y_begin));
            const float um = bload(rub, xo, row_off(y_begin - 1));
            uv[0][0] = (owned & in_rows(y_begin)) ? u0 : 0.0f;
            uv[0][1] = (owned & in_rows(y_begin - 1)) ? um : 0.0f;
        }
        kh[0] = bload(rk, xo, row_off(y_begin));
        kh[1] = bload(rk, xo, row_off(y_begin - 1));
        float qb[kPfA], qa[kPfA], qk[kPfA], qe[kPfA], qm[kPfA], qu[kPfA], qv[kPfA];
        auto fill = [&](int i, int y) {
            qb[i] = bload(rb, xo, row_off(y + 1));
            qa[i] = bload(ra, xo, row_off(y));
            qk[i] = bload(rk, xo, row_off(y + 1));
            qu[i] = bload(rub, xo, row_off(y + 1));
            qv[i] = bload(rua, xo, row_off(y));
            // every iteration issues the same loads, so the compiler's waits
            // count exactly (a conditional load makes them wait for loads
            // issued after it)
            qe[i] = gload_off(emap, xo + row_off(y));
            qm[i] = 0.0f;
            if constexpr (XE) qm[i] = gload_off(mur, mo + mrow(y));
        };
#pragma unroll
        for (int i = 0; i < kPfA; ++i) {
            __builtin_amdgcn_sched_barrier(0);
            fill(i, y_begin + i);
        }
        __builtin_amdgcn_sched_barrier(0);
        if (edge_lane) {
            AdjEdge e;
            e.v[0] = make_float2(kh[0] * lv[0][0], uv[0][0]);
#pragma unroll
            for (int j = 1; j < kMaxLevels; ++j) e.v[j] = make_float2(0.0f, 0.0f);
            ex[y_begin & 1][wave + 1][my_side] = e;
        }
        __syncthreads();
        auto iteration = [&](int y, int i) {
#pragma unroll
            for (int j = 0; j <= KL; ++j) {
                lv[j][2] = lv[j][1];
                lv[j][1] = lv[j][0];
                uv[j][2] = uv[j][1];
                uv[j][1] = uv[j][0];
            }
#pragma unroll
            for (int j = KL + 1; j >= 1; --j) kh[j] = kh[j - 1];
#pragma unroll
            for (int j = KL - 1; j >= 1; --j) {
                mh[j] = mh[j - 1];
                eh[j] = eh[j - 1];
            }
            const float b1 = take(qb[i]), a0 = take(qa[i]), k1 = take(qk[i]), e0 = take(qe[i]),
                        u1 = take(qu[i]), ua = take(qv[i]);
            const float m0 = XE ? take(qm[i]) : 0.0f;
            __builtin_amdgcn_sched_barrier(0);
            fill(i, min(y + kPfA, y_end));
            lv[0][0] = b1;
            uv[0][0] = (owned & in_rows(y + 1)) ? u1 : 0.0f;
            kh[0] = k1;
            eh[0] = owned ? __float_as_int(e0) : -1;
            mh[0] = m0;
#pragma unroll
            for (int j = 1; j <= KL; ++j) {
                const int rho = y - j + 1;
                const int m = n_top - j + 1;  // adjoint step of the level
                // neighbour (P, U) pairs of level j - 1 (one 64-bit LDS read each)
                const float2 ew = ex[y & 1][w_slot][w_side].v[j - 1];
                const float2 ee = ex[y & 1][e_slot][e_side].v[j - 1];
                // adjoint step m (adjm_kernel's arithmetic)
                const float u = lv[j - 1][1];
                const float old = j == 1 ? a0 : lv[j - 2][2];
                const float Ps = kh[j + 1] * lv[j - 1][2];
                const float P0 = kh[j] * u;
                const float Pn = kh[j - 1] * lv[j - 1][0];
                const float pw = from_west(P0, ew.x), pe = from_east(P0, ee.x);
                float v = __fmaf_rn(dt2, lap5(pw, pe, Ps, Pn, P0, p.idx2, p.idy2, p.cen),
                                    leap(u, old));
                // the levels' residual rows lie res_step floats apart
                v = inject_lanes(p, p.lv_resid[0] - (size_t)(j - 1) * res_step, res_off, eh[j - 1],
                                 lane, v);
                if constexpr (XE) {
                    if (!YE || (rho >= 1 && rho <= H - 2)) v = mur_sides(v, u, x, W, mh[j - 1]);
                }
                if constexpr (YE) {
                    if (rho == H - 1) v = mur_row(mur, W, H, x, kt_x, u, lv[j - 1][2], lv[j][1], 2);
                    if (rho == 1) {
                        const float v0 = mur_row(mur, W, H, x, kb_x, lv[j - 1][2], u, v, 0);
                        lv[j][1] = v0;
                        if ((j >= KL - 1) & (r0 == 0)) bstore(ro[j - 1], v0, so, 0);
                    }
                }
                lv[j][0] = v;
                if (j >= KL - 1) {
                    // stores of rows outside the segment are dropped, not
                    // branched around (the same operations every iteration)
                    const bool st = (rho >= r0) & (rho < r1) & (rho > 0);
                    bstore(ro[j - 1], v, st ? so : kDropOffset, st ? rho * row_bytes : 0);
                }
                // the forward field U(m - 2) from U(m - 1) (uv[j - 1]) and U(m)
                const float U0 = uv[j - 1][1];
                const float Uold = j == 1 ? ua : uv[j - 2][2];
                const float Uw = from_west(U0, ew.y), Ue = from_east(U0, ee.y);
                const float Lr = lap5(Uw, Ue, uv[j - 1][2], uv[j - 1][0], U0, p.idx2, p.idy2, p.cen);
                float w = __fmaf_rn(dt2 * kh[j], Lr, leap(U0, Uold));
                if (__builtin_expect((unsigned)(rho - sr) <= (unsigned)src_extra, 0))
                    w = add_transmitter(p, s, w, rho, x, sr, sc, ((const kcf *)p.src_values)[m]);
                const int q = recon_field(n_top, j);
                if constexpr (XE) {
                    const float bc = bcol[((j - 1) * 2 + col_side) * nb + rho - yb];
                    if (side_lane & (rho >= 1) & (rho <= H - 2)) w = bc;
                }
                if constexpr (YE) {
                    if ((rho == 0) | (rho == H - 1)) {  // uniform
                        const int which = rho == 0 ? 0 : brow_second;
                        w = brow[((size_t)which * KL + j - 1) * W + xo / 4];
                    }
                    w = in_rows(rho) ? w : 0.0f;
                }
                if constexpr (XE) w = owned ? w : 0.0f;
                if (q < 0) w = 0.0f;  // U(-1) = 0
                uv[j][0] = w;
                if (j >= KL - 1) {
                    const bool st = (rho >= r0) & (rho < r1);
                    bstore(rou[j - 1], w, st ? so : kDropOffset, st ? rho * row_bytes : 0);
                }
                // imaging of row rho - 1: lam' of step m against Lap(U(m - 1))
                if (EVERY || m >= p.img_from) racc[j - 1] += lv[j][1] * lsave[j - 1];
                lsave[j - 1] = Lr;
            }
            // row y - KL has the terms of every level: add it to the image
            const int Rf = y - KL;
            if ((Rf >= r0) & (Rf < r1) & owned) acc[(Rf - r0) * W + x] += racc[KL - 1];
#pragma unroll
            for (int k = KL - 1; k >= 1; --k) racc[k] = racc[k - 1];
            racc[0] = 0.0f;
            if (edge_lane) {
                AdjEdge e;
#pragma unroll
                for (int j = 0; j < kMaxLevels; ++j)
                    e.v[j] = j < KL ? make_float2(kh[j] * lv[j][0], uv[j][0]) : make_float2(0.0f, 0.0f);
                ex[(y + 1) & 1][wave + 1][my_side] = e;
            }
            __syncthreads();
        };
        // exactly y_end - y_begin + 1 iterations (one barrier each): every
        // wavefront of the workgroup, whichever march it runs, passes the
        // same number of barriers
        static_assert(kPfA == 2, "the march below takes its prefetch slots in pairs");
        for (int y = y_begin;; y += kPfA) {
            iteration(y, 0);
            if (y + 1 > y_end) break;
            iteration(y + 1, 1);
            if (y + 2 > y_end) break;
        }
    }
}

// Lean march of a reconstructing adjoint launch, for a segment whose march
// reaches neither row 1 nor row H - 1 in a launch whose every level images
// and rebuilds a field of step >= 0: SIDE 0 = a wavefront holding neither
// column 0 nor W - 1, SIDE 1 = the one holding column 0, SIDE 2 = the one
// holding column W - 1 (the row spans at least two wavefronts). Per cell it
// does adjr_wave's arithmetic in the same order (the same bits); the
// bookkeeping around it is taken off the march:
// * rings of registers indexed by the iteration count c, made static by
//   unrolling six iterations (a multiple of every ring's period), so no row
//   moves between registers: each level's rows and their P = c^2 lam (three
//   live of six), the level-0 inputs, which the prefetch loads straight into
//   the ring (slot c + 2 holds a row no longer needed), the scatter entries,
//   Mur side coefficients and imaging sums of the rows in flight;
// * P = c^2 lam is formed once, when its row is made, instead of per use;
// * the first 2 KL - 1 iterations (they make rows below the segment, whose
//   values are incomplete) are unrolled with their store tests resolved at
//   compile time. After them every row a level stores or images is the
//   segment's own or -- in the last two iterations -- row r1 or r1 + 1 of
//   levels KL - 1 and KL, whose cells the next segment makes from the same
//   inputs and stores with the same bits, so no test remains;
// * one residual-injection test per iteration for the KL rows in flight;
// * a side wavefront updates only its own Mur column (two wavefront shifts
//   per level instead of four).
template <int KL, int SIDE>
__device__ __forceinline__ void adjr_lean(const StepParams &p, int lane, int wave, int grp,
                                          int seg, AdjEdgeRow *ex, float *bcol)
{
    constexpr int kR = 6;             // ring length = iterations per unrolled group
    constexpr int kPro = 2 * KL - 1;  // iterations before every stored row is owned
    constexpr int kPf = kPfA;         // rows loaded ahead
    static_assert(kR % kPf == 0 && kPf + KL <= kR, "prefetched rows must fit the rings");
    const int W = p.nx, H = p.ny;
    const int x = wave * kLanes - p.wide_off + lane;
    const bool owned = SIDE == 0 || ((x >= 0) & (x < W));
    const int xo = (SIDE == 0 ? x : min(max(x, 0), W - 1)) * 4;
    const int so = owned ? x * 4 : kDropOffset;
    // the Mur column of a side wavefront, its side coefficients (left ones
    // for column 0, right ones for W - 1) and its staged boundary column
    const bool side_lane = SIDE == 1 ? x == 0 : (SIDE == 2 ? x == W - 1 : false);
    const int mo = SIDE == 2 ? H * 4 : 0;
    const float *mur = p.mur;
    int r0, r1;
    seg_rows(p, seg, r0, r1);
    const int y_begin = r0 - (KL - 1);
    const int y_end = r1 + KL - 1;
    const int n_it = y_end - y_begin + 1;
    const size_t plane = (size_t)H * W;
    const unsigned plane_bytes = (unsigned)(plane * sizeof(float));
    const int row_bytes = W * 4;
    const __amdgpu_buffer_rsrc_t rk = make_rsrc(p.c2, plane_bytes);
    const float *emap = (const float *)p.cell_map;
    const float dt2 = p.dt2, idx2 = p.idx2, idy2 = p.idy2, cen = p.cen;
    const int w_slot = lane == 0 ? wave : wave + 1;
    const int w_side = lane == 0 ? 1 : 0;
    const int e_slot = lane == kLanes - 1 ? wave + 2 : wave + 1;
    const int e_side = lane == kLanes - 1 ? 0 : 1;
    const bool edge_lane = (lane == 0) | (lane == kLanes - 1);
    const int my_side = lane == kLanes - 1 ? 1 : 0;
    const int n_top = p.src_index;
    // staged boundary column of level j at row rho: bcol[((j - 1) * 2 + side) *
    // nb + rho - yb] with yb = y_begin - KL + 1, i.e. rho - yb = c + KL - j
    const int nb = p.bcol_rows;
    const float *bside = bcol + (SIDE == 2 ? nb : 0) + KL;
    // edge rows by iteration parity: iteration c reads ex_c[c & 1] and writes
    // ex_c[(c + 1) & 1] (the rings' length is even, so c & 1 is static)
    AdjEdgeRow *const ex_c[2] = {ex + (y_begin & 1), ex + ((y_begin + 1) & 1)};
    const int s_begin = grp * p.shots_per_group;
    const int s_end = min(p.n_shots, s_begin + p.shots_per_group);
    const int src_extra = transmitter_rows(p);
    const size_t res_step = (size_t)p.n_shots * p.n_rec;
    for (int s = s_begin; s < s_end; ++s) {
        adjr_stage_columns<KL>(p, s, y_begin - KL + 1, bcol);
        const size_t soff = (size_t)s * plane;
        const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.f_old + soff, plane_bytes);
        const __amdgpu_buffer_rsrc_t rb = make_rsrc(p.f_cur + soff, plane_bytes);
        const __amdgpu_buffer_rsrc_t rua = make_rsrc(p.r_next + soff, plane_bytes);
        const __amdgpu_buffer_rsrc_t rub = make_rsrc(p.r_cur + soff, plane_bytes);
        __amdgpu_buffer_rsrc_t ro[KL], rou[KL];
#pragma unroll
        for (int j = 0; j < KL; ++j) {
            ro[j] = make_rsrc((const float *)((uintptr_t)p.lv_out[j] + soff * sizeof(float)),
                              plane_bytes);
            rou[j] = make_rsrc((const float *)((uintptr_t)p.lv_rout[j] + soff * sizeof(float)),
                               plane_bytes);
        }
        const size_t res_off = (size_t)s * p.n_rec;
        const int sr = p.src_rows[s], sc = p.src_cols[s];
        // ring slot of iteration c = c mod kR. Level 0 (the inputs): lam[0],
        // U[0] = row y + 1 at iteration c; lamo, uno = the older adjoint and
        // the newer forward field at row y; c2 = c^2 of row y + 1; ent, mh =
        // the scatter entry and Mur side coefficient at row y. Level j >= 1:
        // lam[j], U[j] = the row made at iteration c (row y - j + 1); P[j] =
        // c^2 lam of the same rows. racc: imaging sum of row y_begin + slot
        // (mod kR). The rows in use plus those in flight fit every ring but
        // that of c^2 (in use for KL + 1 iterations), which the prefetch
        // fills through slots qk.
        float lam[KL + 1][kR], P[KL][kR], U[KL][kR];
        float c2[kR], lamo[kR], uno[kR], racc[kR], mh[kR], lsave[KL], qk[kPf];
        int ent[kR];
#pragma unroll
        for (int k = 0; k < kR; ++k) {
#pragma unroll
            for (int j = 0; j <= KL; ++j) lam[j][k] = 0.0f;
#pragma unroll
            for (int j = 0; j < KL; ++j) P[j][k] = U[j][k] = 0.0f;
            c2[k] = lamo[k] = uno[k] = racc[k] = mh[k] = 0.0f;
            ent[k] = -1;
        }
#pragma unroll
        for (int j = 0; j < KL; ++j) lsave[j] = 0.0f;
        // iterations -1 and -2: rows y_begin and y_begin - 1
        lam[0][kR - 1] = bload(rb, xo, y_begin * row_bytes);
        lam[0][kR - 2] = bload(rb, xo, (y_begin - 1) * row_bytes);
        U[0][kR - 1] = bload(rub, xo, y_begin * row_bytes);
        U[0][kR - 2] = bload(rub, xo, (y_begin - 1) * row_bytes);
        c2[kR - 1] = bload(rk, xo, y_begin * row_bytes);
        c2[kR - 2] = bload(rk, xo, (y_begin - 1) * row_bytes);
#pragma unroll
        for (int i = 0; i < kPf; ++i) {
            __builtin_amdgcn_sched_barrier(0);
            const int yy = min(y_begin + i, y_end);
            const int o0 = yy * row_bytes, o1 = o0 + row_bytes;
            lam[0][i] = bload(rb, xo, o1);
            U[0][i] = bload(rub, xo, o1);
            qk[i] = bload(rk, xo, o1);
            lamo[i] = bload(ra, xo, o0);
            uno[i] = bload(rua, xo, o0);
            ent[i] = __float_as_int(gload_off(emap, xo + o0));
            if constexpr (SIDE != 0) mh[i] = gload_off(mur, mo + yy * 4);
        }
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (SIDE != 0) {
            U[0][kR - 1] = owned ? U[0][kR - 1] : 0.0f;
            U[0][kR - 2] = owned ? U[0][kR - 2] : 0.0f;
        }
        P[0][kR - 1] = c2[kR - 1] * lam[0][kR - 1];
        P[0][kR - 2] = c2[kR - 2] * lam[0][kR - 2];
        if (edge_lane) {
            AdjEdge e;
            e.v[0] = make_float2(P[0][kR - 1], U[0][kR - 1]);
#pragma unroll
            for (int j = 1; j < kMaxLevels; ++j) e.v[j] = make_float2(0.0f, 0.0f);
            (*ex_c[0])[wave + 1][my_side] = e;
        }
        __syncthreads();
        // iteration c with ring slot cm = c mod kR; pc = c in the first kPro
        // iterations (row tests at compile time), -1 afterwards (no tests)
        auto iteration = [&](int cm, int pc, int c) __attribute__((always_inline)) {
            const int y = y_begin + c;
            const int cn = cm, cc = (cm + kR - 1) % kR, cs = (cm + kR - 2) % kR;
            const int q = cm % kPf;
            const float k_new = take(qk[q]);
            __builtin_amdgcn_sched_barrier(0);
            {
                const int f = (cm + kPf) % kR;
                const int yp = min(y + kPf, y_end);
                const int o0 = yp * row_bytes, o1 = o0 + row_bytes;
                lam[0][f] = bload(rb, xo, o1);
                U[0][f] = bload(rub, xo, o1);
                qk[q] = bload(rk, xo, o1);
                lamo[f] = bload(ra, xo, o0);
                uno[f] = bload(rua, xo, o0);
                ent[f] = __float_as_int(gload_off(emap, xo + o0));
                if constexpr (SIDE != 0) mh[f] = gload_off(mur, mo + yp * 4);
            }
            __builtin_amdgcn_sched_barrier(0);
            c2[cm] = k_new;
            P[0][cm] = k_new * lam[0][cm];
            if constexpr (SIDE != 0) U[0][cm] = owned ? U[0][cm] : 0.0f;
            // a lane's entries of rows y .. y - KL + 1 have an AND >= 0 iff
            // one of them is >= 0
            int e_and = ent[cm];
#pragma unroll
            for (int j = 2; j <= KL; ++j) e_and &= ent[(cm + kR - (j - 1)) % kR];
            const bool inject = __ballot(owned & (e_and >= 0)) != 0;
            const bool near_src = (unsigned)(y - sr) <= (unsigned)(KL - 1 + src_extra);
            const AdjEdgeRow &er = *ex_c[cm & 1];
#pragma unroll
            for (int j = 1; j <= KL; ++j) {
                const int rho = y - j + 1;
                const int kj = (cm + kR - j) % kR;         // c^2 of row rho
                const int ke = (cm + kR - (j - 1)) % kR;   // entry, Mur coefficient of row rho
                const float2 ew = er[w_slot][w_side].v[j - 1];
                const float2 ee = er[e_slot][e_side].v[j - 1];
                // adjoint step n_top - j + 1
                const float Pn = P[j - 1][cn], P0 = P[j - 1][cc], Ps = P[j - 1][cs];
                const float u = lam[j - 1][cc];
                const float old = j == 1 ? lamo[cm] : lam[j - 2][cs];
                const float pw = from_west(P0, ew.x), pe = from_east(P0, ee.x);
                float v = __fmaf_rn(dt2, lap5(pw, pe, Ps, Pn, P0, idx2, idy2, cen),
                                    leap(u, old));
                if (inject)
                    v = inject_lanes(p, p.lv_resid[0] - (size_t)(j - 1) * res_step, res_off,
                                     owned ? ent[ke] : -1, lane, v);
                if constexpr (SIDE == 1) {  // column 0: from column 1's new and current values
                    const float ve = from_east(v, 0.0f), ce = from_east(u, 0.0f);
                    v = side_lane ? ce + mh[ke] * (ve - u) : v;
                } else if constexpr (SIDE == 2) {  // column W - 1: from column W - 2's
                    const float vw = from_west(v, 0.0f), cw = from_west(u, 0.0f);
                    v = side_lane ? cw + mh[ke] * (vw - u) : v;
                }
                lam[j][cm] = v;
                if (j < KL) P[j][cm] = c2[kj] * v;
                const bool store = j >= KL - 1 && (pc < 0 || pc >= KL + j - 2);
                if (store) bstore(ro[j - 1], v, so, rho * row_bytes);
                // the forward field U(n_top - j - 1) from U(n_top - j) and U(n_top - j + 1)
                const float U0 = U[j - 1][cc];
                const float Uold = j == 1 ? uno[cm] : U[j - 2][cs];
                const float Uw = from_west(U0, ew.y), Ue = from_east(U0, ee.y);
                const float Lr = lap5(Uw, Ue, U[j - 1][cs], U[j - 1][cn], U0, idx2, idy2, cen);
                float w = __fmaf_rn(dt2 * c2[kj], Lr, leap(U0, Uold));
                if (__builtin_expect(near_src, 0) &&
                    (unsigned)(rho - sr) <= (unsigned)src_extra)
                    w = add_transmitter(p, s, w, rho, x, sr, sc,
                                        ((const kcf *)p.src_values)[n_top - j + 1]);
                if constexpr (SIDE != 0) {
                    // the Mur column from the boundary ring; no field outside the grid
                    const float bc = bside[(j - 1) * 2 * nb + c - j];
                    w = side_lane ? bc : w;
                    w = owned ? w : 0.0f;
                }
                if (j < KL) U[j][cm] = w;
                if (store) bstore(rou[j - 1], w, so, rho * row_bytes);
                // imaging of row rho - 1: lam' of the step against Lap(U) of the next
                racc[kj] += lam[j][cc] * lsave[j - 1];
                lsave[j - 1] = Lr;
            }
            // row y - KL has the terms of every level (rows below the
            // segment, in the first kPro iterations, are dropped)
            const int kf = (cm + kR - KL) % kR;
            if (pc < 0 && owned) adjm_acc[(c - kPro) * W + x] += racc[kf];
            racc[kf] = 0.0f;
            if (edge_lane) {
                AdjEdge e;
#pragma unroll
                for (int j = 0; j < kMaxLevels; ++j)
                    e.v[j] = j < KL ? make_float2(P[j][cm], U[j][cm]) : make_float2(0.0f, 0.0f);
                (*ex_c[(cm + 1) & 1])[wave + 1][my_side] = e;
            }
            __syncthreads();
        };
#pragma unroll
        for (int c = 0; c < kPro; ++c) iteration(c % kR, c, c);
        // the rest in groups of kR from ring slot kPro mod kR; exactly n_it
        // iterations in all (one barrier each, as in adjr_wave)
        int c = kPro;
        for (;;) {
#pragma unroll
            for (int k = 0; k < kR; ++k) {
                iteration((kPro + k) % kR, -1, c);
                if (++c == n_it) goto march_done;
            }
        }
    march_done:;
    }
}

// Workgroup b (row segment, shot group) of a reconstructing adjoint launch.
template <int KL, bool EVERY>
__device__ __forceinline__ void adjr_block(const StepParams &p, int b, AdjEdgeRow *ex)
{
    const int wblk = (b % kXcds) * p.chunk + b / kXcds;
    if (wblk >= p.n_work) return;  // uniform over the workgroup
    const int t = (int)threadIdx.x;
    if (t < 2 * 2 * kMaxLevels) {  // zero pads at both ends, both parities
        const int par = t / (2 * kMaxLevels), side = (t / kMaxLevels) & 1, j = t % kMaxLevels;
        ex[par][0][side].v[j] = make_float2(0.0f, 0.0f);
        ex[par][p.wide_waves + 1][side].v[j] = make_float2(0.0f, 0.0f);
    }
    const int lane = t & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane(t / kLanes);
    const int grp = wblk % p.n_groups;
    const int seg = wblk / p.n_groups;
    int r0, r1;
    seg_rows(p, seg, r0, r1);
    const int W = p.nx;
    const int n_acc = (r1 - r0) * W;
    for (int i = t; i < n_acc; i += (int)blockDim.x) adjm_acc[i] = 0.0f;
    float *bcol = adjm_acc + (size_t)p.acc_rows * W;
    float *brow = bcol + (size_t)KL * 2 * p.bcol_rows;
    // (the first barrier of the march orders these writes before any use)
    const bool xe = (wave == 0) || (wave == p.wide_waves - 1);
    const bool ye = (r0 < 2 * KL) || (r1 + KL >= p.ny);
    // lean marches (segments away from rows 0 / H - 1 of launches whose
    // every level images and rebuilds a field; rows of two or more
    // wavefronts) and the general one, with the same bits
    bool lean = false;
    if constexpr (EVERY) lean = !ye && p.adjr_lean;
    if (lean) {
        if (wave == 0) adjr_lean<KL, 1>(p, lane, wave, grp, seg, ex, bcol);
        else if (wave == p.wide_waves - 1) adjr_lean<KL, 2>(p, lane, wave, grp, seg, ex, bcol);
        else adjr_lean<KL, 0>(p, lane, wave, grp, seg, ex, bcol);
    } else if (xe) {
        if (ye) adjr_wave<KL, true, true, EVERY>(p, lane, wave, grp, seg, ex, bcol, brow);
        else adjr_wave<KL, true, false, EVERY>(p, lane, wave, grp, seg, ex, bcol, brow);
    } else {
        if (ye) adjr_wave<KL, false, true, EVERY>(p, lane, wave, grp, seg, ex, bcol, brow);
        else adjr_wave<KL, false, false, EVERY>(p, lane, wave, grp, seg, ex, bcol, brow);
    }
    __syncthreads();
    // partial[grp] += scale * segment image, per cell
    const int x = wave * kLanes - p.wide_off + lane;
    if ((x >= 0) & (x < W)) {
        gf *gp = gstore_ptr(p.g_part) + (size_t)grp * ((size_t)p.ny * W);
        add_image_column(gp + (size_t)r0 * W + x, r1 - r0, W, p.img_scale, adjm_acc + x, W);
    }
}

template <int KL, bool EVERY>
__global__ void __launch_bounds__(1024) adjr_kernel(StepParams p)
{
    __shared__ AdjEdgeRow ex[2];
    adjr_block<KL, EVERY>(p, (int)blockIdx.x, ex);
}

// Batched reconstructing adjoint launch (entries of one row width; the LDS is
// sized for the largest entry).
template <int KL, bool EVERY>
__global__ void __launch_bounds__(1024) adjr_batch_kernel(const StepParams *table,
                                                          const int *first, int n)
{
    __shared__ AdjEdgeRow ex[2];
    const int b = (int)blockIdx.x;
    const int e = __builtin_amdgcn_readfirstlane(batch_entry(first, n, b));
    adjr_block<KL, EVERY>(batch_params(table, e), b - first[e], ex);
}

// g[i] += sum over groups of g_part[k][i], in group order.
__global__ void __launch_bounds__(256) reduce_groups_kernel(size_t plane, int n_groups,
                                                            const float *g_part, float *g)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= plane) return;
    float acc = 0.0f;
    for (int k = 0; k < n_groups; ++k) acc += g_part[(size_t)k * plane + i];
    g[i] += acc;
}

__global__ void __launch_bounds__(kGatherThreads) gather_kernel(StepParams p)
{
    gather_pairs(p, blockIdx.x * (int)blockDim.x + (int)threadIdx.x);
}

// Bilinear gather of every level of a multi-step launch (p.levels fields
// p.lv_out[j] into p.lv_samples[j]), same arithmetic as gather_pairs.
__global__ void __launch_bounds__(kGatherThreads) gather_levels_kernel(StepParams p)
{
    const int per = p.n_shots * p.n_rec;
    int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = t / per;
    if (j >= p.levels) return;
    t -= j * per;
    const float *field = p.lv_out[j];
    float *dest = p.lv_samples[j];
    const int R = p.n_rec;
    const int s = t / R;
    const int r = t - s * R;
    const int W = p.nx;
    const float *f = field + (size_t)s * ((size_t)p.ny * W);
    const size_t i = (size_t)p.rec_row0[r] * W + p.rec_col0[r];
    const float *w = p.rec_w;
    float acc = f[i] * w[r];
    acc = acc + f[i + 1] * w[R + r];
    acc = acc + f[i + W] * w[2 * R + r];
    acc = acc + f[i + W + 1] * w[3 * R + r];
    dest[t] = acc;
}

// Batched form of gather_levels_kernel: entry e of `table` is one run's
// launch (its levels, fields and receivers); its workgroups are numbered from
// first[e].
__global__ void __launch_bounds__(kGatherThreads) gather_levels_batch_kernel(
    const StepParams *table, const int *first, int n)
{
    const int b = (int)blockIdx.x;
    const int e = __builtin_amdgcn_readfirstlane(batch_entry(first, n, b));
    const StepParams &p = batch_params(table, e);
    const int per = p.n_shots * p.n_rec;
    int t = (b - first[e]) * (int)blockDim.x + (int)threadIdx.x;
    const int j = t / per;
    if (j >= p.levels) return;
    t -= j * per;
    const float *field = p.lv_out[j];
    float *dest = p.lv_samples[j];
    const int R = p.n_rec;
    const int s = t / R;
    const int r = t - s * R;
    const int W = p.nx;
    const float *f = field + (size_t)s * ((size_t)p.ny * W);
    const size_t i = (size_t)p.rec_row0[r] * W + p.rec_col0[r];
    const float *w = p.rec_w;
    float acc = f[i] * w[r];
    acc = acc + f[i + 1] * w[R + r];
    acc = acc + f[i + W] * w[2 * R + r];
    acc = acc + f[i + W + 1] * w[3 * R + r];
    dest[t] = acc;
}

// Stand-alone imaging condition: g_out = g_in + img_scale * sum_s f_cur[s] *
// Lap(img[s]), with the same per-cell accumulation order as the fused path.
__global__ void __launch_bounds__(256) image_kernel(StepParams p)
{
    const int W = p.nx, H = p.ny;
    const size_t plane = (size_t)H * W;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= plane) return;
    const int y = (int)(i / W);
    const int x = (int)(i - (size_t)y * W);
    float acc = 0.0f;
    for (int s = 0; s < p.n_shots; ++s) {
        const size_t soff = (size_t)s * plane;
        const float *f = p.img + soff;
        const float fw = x > 0 ? f[i - 1] : 0.0f;
        const float fe = x < W - 1 ? f[i + 1] : 0.0f;
        const float fs = y > 0 ? f[i - W] : 0.0f;
        const float fn = y < H - 1 ? f[i + W] : 0.0f;
        // a Laplacian history holds lf itself
        const float lf = p.hist_lap ? f[i] : lap5(fw, fe, fs, fn, f[i], p.idx2, p.idy2, p.cen);
        acc = __fmaf_rn(p.f_cur[soff + i], lf, acc);
    }
    p.g_out[i] = p.g_in[i] + p.img_scale * acc;
}

__global__ void prepare_medium_kernel(int ny, int nx, const float *c, float dt, float dx,
                                      float dy, float dt2, float *c2, float *c2dt2,
                                      float *mur)
{
    const size_t n = (size_t)ny * nx;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float cv = c[i];
        const float sq = cv * cv;
        c2[i] = sq;
        c2dt2[i] = dt2 * sq;
    }
    if (i < (size_t)ny) {
        const float cl = c[i * nx];
        const float cr = c[i * nx + nx - 1];
        mur[i] = (cl * dt - dx) / (cl * dt + dx);
        mur[ny + i] = (cr * dt - dx) / (cr * dt + dx);
    }
    if (i < (size_t)nx) {
        const float cb = c[i];
        const float ct = c[(size_t)(ny - 1) * nx + i];
        mur[2 * ny + i] = (cb * dt - dy) / (cb * dt + dy);
        mur[2 * ny + nx + i] = (ct * dt - dy) / (ct * dt + dy);
    }
    if (i < 4) {
        const size_t r = (i < 2) ? 0 : (size_t)(ny - 1);
        const size_t cc = (i % 2 == 0) ? 0 : (size_t)(nx - 1);
        const float cv = c[r * nx + cc];
        mur[2 * ny + 2 * nx + i] = (cv * dt - dx) / (cv * dt + dx);
        mur[2 * ny + 2 * nx + 4 + i] = (cv * dt - dy) / (cv * dt + dy);
    }
}

__global__ void fill_kernel(float *p, size_t n, float v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Boundary cells of `field` (S, H, W) into a slot of the boundary ring (S,
// bnd_floats): after launches that store their fields in full.
__device__ __forceinline__ void bnd_extract_cell(const float *field, int S, int H, int W,
                                                 float *slot, size_t i)
{
    const size_t P = bnd_floats(H, W);
    if (i >= (size_t)S * P) return;
    const size_t s = i / P;
    const int k = (int)(i % P);
    int row, col;
    if (k < W) {
        row = 0;
        col = k;
    } else if (k < 2 * W) {
        row = H - 1;
        col = k - W;
    } else if (k < 2 * W + H - 2) {
        row = k - 2 * W + 1;
        col = 0;
    } else {
        row = k - 2 * W - (H - 2) + 1;
        col = W - 1;
    }
    slot[i] = field[s * H * W + (size_t)row * W + col];
}

// The fields of a launch's levels and their ring slots, for one extraction
// launch over all of them (blockIdx.y = level).
struct BndLevels {
    const float *field[kMaxLevels];
    float *slot[kMaxLevels];
};

__global__ void __launch_bounds__(256) bnd_extract_kernel(BndLevels lv, int S, int H, int W)
{
    const int j = (int)blockIdx.y;
    bnd_extract_cell(lv.field[j], S, H, W, lv.slot[j],
                     (size_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// Batched ring extraction: entry e copies the boundary cells of its n_shots
// fields at p.f_cur into the ring slot at p.bnd.
__global__ void __launch_bounds__(256) bnd_extract_batch_kernel(const StepParams *table,
                                                                const int *first, int n)
{
    const int b = (int)blockIdx.x;
    const int e = __builtin_amdgcn_readfirstlane(batch_entry(first, n, b));
    const StepParams &p = batch_params(table, e);
    bnd_extract_cell(p.f_cur, p.n_shots, p.ny, p.nx, p.bnd,
                     (size_t)(b - first[e]) * blockDim.x + threadIdx.x);
}

// Reduced-precision history storage: float32 <-> 16-bit codes, both rounded
// to nearest even. float16 goes through the hardware conversion
// (v_cvt_f16_f32, IEEE binary16 with subnormals); bfloat16 keeps the upper
// half of the float32 bits after adding the round-to-nearest-even bias.
__device__ __forceinline__ uint16_t f32_to_code(float v, int type)
{
    if (type == SWFD_HISTORY_F16) {
        const _Float16 h = (_Float16)v;
        return __builtin_bit_cast(uint16_t, h);
    }
    const uint32_t b = __float_as_uint(v);
    if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((b >> 16) | 0x40u);  // quiet NaN
    return (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float code_to_f32(uint16_t c, int type)
{
    if (type == SWFD_HISTORY_F16) return (float)__builtin_bit_cast(_Float16, c);
    return __uint_as_float((uint32_t)c << 16);
}

// Eight values per lane when every pointer is 16-byte aligned and n % 8 == 0
// (two float4 <-> one uint4), else one per lane.
template <bool kPack, int kVec>
__global__ void __launch_bounds__(256) history_convert_kernel(size_t n, const void *src,
                                                              void *dst, int type)
{
    const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * kVec;
    if (i >= n) return;
    if constexpr (kVec == 8) {
        if constexpr (kPack) {
            const float4 a = reinterpret_cast<const float4 *>(src)[i / 4];
            const float4 b = reinterpret_cast<const float4 *>(src)[i / 4 + 1];
            const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
            uint32_t w[4];
            for (int k = 0; k < 4; ++k)
                w[k] = (uint32_t)f32_to_code(v[2 * k], type) |
                       ((uint32_t)f32_to_code(v[2 * k + 1], type) << 16);
            reinterpret_cast<uint4 *>(dst)[i / 8] = make_uint4(w[0], w[1], w[2], w[3]);
        } else {
            const uint4 c = reinterpret_cast<const uint4 *>(src)[i / 8];
            const uint32_t w[4] = {c.x, c.y, c.z, c.w};
            float v[8];
            for (int k = 0; k < 4; ++k) {
                v[2 * k] = code_to_f32((uint16_t)(w[k] & 0xffffu), type);
                v[2 * k + 1] = code_to_f32((uint16_t)(w[k] >> 16), type);
            }
            reinterpret_cast<float4 *>(dst)[i / 4] = make_float4(v[0], v[1], v[2], v[3]);
            reinterpret_cast<float4 *>(dst)[i / 4 + 1] = make_float4(v[4], v[5], v[6], v[7]);
        }
    } else {
        if constexpr (kPack)
            static_cast<uint16_t *>(dst)[i] = f32_to_code(static_cast<const float *>(src)[i], type);
        else
            static_cast<float *>(dst)[i] = code_to_f32(static_cast<const uint16_t *>(src)[i], type);
    }
}

__global__ void scatter_kernel(int n_shots, size_t plane, int n_entries, int n_rec,
                               const float *resid, const int *entry_cell,
                               const int *entry_start, const int *pair_rec,
                               const float *pair_w, float *out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_shots * n_entries) return;
    const int s = t / n_entries;
    const int e = t - s * n_entries;
    const float *res = resid + (size_t)s * n_rec;
    float acc = 0.0f;
    for (int k = entry_start[e]; k < entry_start[e + 1]; ++k)
        acc += res[pair_rec[k]] * pair_w[k];
    out[(size_t)s * plane + entry_cell[e]] = acc;
}

__global__ void laplacian_kernel(int ny, int nx, size_t total, float idx2, float idy2,
                                 float cen, const float *f, float *out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const size_t plane = (size_t)ny * nx;
    const size_t within = i % plane;
    const int y = (int)(within / nx);
    const int x = (int)(within - (size_t)y * nx);
    const float uw = x > 0 ? f[i - 1] : 0.0f;
    const float ue = x < nx - 1 ? f[i + 1] : 0.0f;
    const float us = y > 0 ? f[i - nx] : 0.0f;
    const float un = y < ny - 1 ? f[i + nx] : 0.0f;
    out[i] = lap5(uw, ue, us, un, f[i], idx2, idy2, cen);
}

// residual[n, s, r] = samples[n, s, r] - observed[obs_index[s], r, n] (muted
// channel zeroed) through a 64x64 LDS transpose tile; per-block partial sums
// of residual^2 in a fixed tree order.
constexpr int kResTile = 64;
__global__ void __launch_bounds__(256) residual_kernel(int N, int S, int R,
                                                       const float *samples,
                                                       const float *obs, long long se,
                                                       long long sr, const int *obs_index,
                                                       const int *mute, float *resid,
                                                       float *partials)
{
    __shared__ float tile[kResTile][kResTile + 1];
    __shared__ float red[256];
    const int tiles_r = (R + kResTile - 1) / kResTile;
    const int tiles_n = (N + kResTile - 1) / kResTile;
    int b = blockIdx.x;
    const int s = b / (tiles_r * tiles_n);
    b -= s * tiles_r * tiles_n;
    const int tn = b / tiles_r;
    const int tr = b - tn * tiles_r;
    const int r0 = tr * kResTile, n0 = tn * kResTile;
    const int tx = threadIdx.x & 63, tyy = threadIdx.x >> 6;
    const float *o = obs + (long long)obs_index[s] * se;
    for (int rl = tyy; rl < kResTile; rl += 4) {
        const int r = r0 + rl, n = n0 + tx;
        tile[rl][tx] = (r < R && n < N) ? o[(long long)r * sr + n] : 0.0f;
    }
    __syncthreads();
    const int m = mute[s];
    float local = 0.0f;
    for (int nl = tyy; nl < kResTile; nl += 4) {
        const int n = n0 + nl, r = r0 + tx;
        if (n < N && r < R) {
            const size_t idx = ((size_t)n * S + s) * R + r;
            float d = samples[idx] - tile[tx][nl];
            if (r == m) d = 0.0f;
            resid[idx] = d;
            local += d * d;
        }
    }
    red[threadIdx.x] = local;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// Compensated (Kahan) sum of the partials by one workgroup: each lane sums a
// fixed stride, then lane 0 adds the lane sums in order.
__global__ void __launch_bounds__(256) misfit_kernel(int n_partials, const float *partials,
                                                     float *misfit)
{
    __shared__ float lane_sum[256];
    __shared__ float lane_err[256];
    float sum = 0.0f, comp = 0.0f;
    for (int i = threadIdx.x; i < n_partials; i += 256) {
        const float yv = partials[i] - comp;
        const float t = sum + yv;
        comp = (t - sum) - yv;
        sum = t;
    }
    lane_sum[threadIdx.x] = sum;
    lane_err[threadIdx.x] = comp;
    __syncthreads();
    if (threadIdx.x == 0) {
        float total = 0.0f, c = 0.0f;
        for (int i = 0; i < 256; ++i) {
            const float yv = (lane_sum[i] - lane_err[i]) - c;
            const float t = total + yv;
            c = (t - total) - yv;
            total = t;
        }
        misfit[0] = 0.5f * total;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct LaunchShape {
    int tiles_x, strips, n_groups, shots_per_group;
};

int g_cu_count = 0;

int cu_count()
{
    if (g_cu_count == 0) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
            g_cu_count = prop.multiProcessorCount;
        else
            g_cu_count = 256;
    }
    return g_cu_count;
}

// Work items of an (ny, nx, S) step: 64-column groups x kRows-row strips x
// shot groups. Shots are split until the launch holds about 32 wavefronts per
// CU; with imaging at most kMaxImageGroups groups (one partial each).
int env_int(const char *name, int fallback)
{
    const char *v = std::getenv(name);
    return (v != nullptr && *v != '\0') ? std::atoi(v) : fallback;
}

// Launch-shape overrides for tuning sweeps (0 = automatic):
// SWFD_SHOTS_PER_GROUP fixes the shots per wavefront work item,
// SWFD_WAVES_PER_CU sets the target work items per CU.
LaunchShape choose_shape(int ny, int nx, int n_shots, bool image)
{
    static const int spg_override = env_int("SWFD_SHOTS_PER_GROUP", 0);
    static const int waves_per_cu = env_int("SWFD_WAVES_PER_CU", 128);
    LaunchShape sh{};
    sh.tiles_x = (nx + kLanes - 1) / kLanes;
    sh.strips = (ny + kRows - 1) / kRows;
    const long cells = (long)sh.tiles_x * sh.strips;
    const long target = (long)waves_per_cu * cu_count();
    long groups = (target + cells - 1) / cells;
    if (spg_override > 0) groups = (n_shots + spg_override - 1) / spg_override;
    groups = std::max<long>(1, std::min<long>(groups, n_shots));
    if (image) groups = std::min<long>(groups, kMaxImageGroups);
    sh.shots_per_group = (int)((n_shots + groups - 1) / groups);
    sh.n_groups = (n_shots + sh.shots_per_group - 1) / sh.shots_per_group;
    return sh;
}

void fill_geometry(StepParams &p, const LaunchShape &sh)
{
    p.tiles_x = sh.tiles_x;
    p.n_groups = sh.n_groups;
    p.shots_per_group = sh.shots_per_group;
    p.waves = kWavesPerBlock;
    p.rows = kRows;
    static const int order = env_int("SWFD_ORDER", 1);
    p.order = order;
    p.n_work = sh.tiles_x * sh.strips * sh.n_groups;
    const int blocks = (p.n_work + kWavesPerBlock - 1) / kWavesPerBlock;
    p.chunk = (blocks + kXcds - 1) / kXcds;
    p.n_stencil_blocks = p.chunk * kXcds;
}

void base_params(StepParams &p, const swfd_grid *g, const swfd_medium *m, int n_shots)
{
    std::memset(&p, 0, sizeof p);
    p.ny = g->ny;
    p.nx = g->nx;
    p.n_shots = n_shots;
    p.idx2 = g->inv_dx2;
    p.idy2 = g->inv_dy2;
    p.cen = g->center;
    p.dt2 = g->dt2;
    if (m != nullptr) {
        p.c2 = m->c2;
        p.c2dt2 = m->c2dt2;
        p.mur = m->mur;
    }
}

void attach_receivers(StepParams &p, const swfd_receivers *rec)
{
    if (rec == nullptr) return;
    p.n_rec = rec->n;
    p.rec_row0 = rec->row0;
    p.rec_col0 = rec->col0;
    p.rec_w = rec->weights;
    p.cell_map = rec->cell_map;
    p.entry_start = rec->entry_start;
    p.pair_rec = rec->pair_rec;
    p.pair_w = rec->pair_w;
}

unsigned step_flags(const StepParams &p, bool gather)
{
    unsigned f = 0;
    if (p.src_rows != nullptr) f |= kPoint;
    if (p.dense != nullptr) f |= kDense;
    if (p.resid != nullptr) f |= kResid;
    if (p.img != nullptr) f |= kImage;
    if (p.f_out2 != nullptr) f |= kOut2;
    if (gather) f |= kGather;
    return f;
}

template <bool ADJ, int RB>
bool launch_variant(unsigned f, dim3 grid, dim3 block, hipStream_t st, const StepParams &p)
{
#define SWFD_CASE(FL)                                                               \
    case (FL):                                                                      \
        hipLaunchKernelGGL((step_kernel<ADJ, RB, (FL)>), grid, block, 0, st, p);  \
        return true;
    if constexpr (!ADJ) {
        switch (f) {
            SWFD_CASE(0u)
            SWFD_CASE(kPoint)
            SWFD_CASE(kPoint | kGather)
            SWFD_CASE(kPoint | kOut2)
            SWFD_CASE(kPoint | kOut2 | kGather)
            SWFD_CASE(kDense)
            SWFD_CASE(kPoint | kDense)
            default:
                return false;
        }
    } else {
        switch (f) {
            SWFD_CASE(0u)
            SWFD_CASE(kResid)
            SWFD_CASE(kResid | kImage)
            SWFD_CASE(kDense)
            SWFD_CASE(kDense | kImage)
            SWFD_CASE(kImage)
            SWFD_CASE(kPoint)
            SWFD_CASE(kPoint | kDense)
            default:
                return false;
        }
    }
#undef SWFD_CASE
}

int launch_step(StepParams &p, bool adjoint, hipStream_t stream, int gather_pairs_count)
{
    const bool image = adjoint && p.img != nullptr;
    const LaunchShape sh = choose_shape(p.ny, p.nx, p.n_shots, image);
    fill_geometry(p, sh);
    const int gather_blocks = gather_pairs_count > 0 ? (gather_pairs_count + 255) / 256 : 0;
    const dim3 block(kLanes * kWavesPerBlock);
    const long edge_items = (long)(2 * (p.nx + p.ny) - 4) * p.n_shots;
    p.n_edge_blocks = (int)((edge_items + block.x - 1) / block.x);
    const dim3 grid(p.n_edge_blocks + p.n_stencil_blocks + gather_blocks);
    const unsigned f = step_flags(p, gather_blocks > 0);
    const bool ok = adjoint ? launch_variant<true, kRows>(f, grid, block, stream, p)
                            : launch_variant<false, kRows>(f, grid, block, stream, p);
    if (!ok) return fail(SWFD_EINVAL, "unsupported combination of step sources / outputs");
    SWFD_HIP_CHECK(hipGetLastError());
    return SWFD_OK;
}

// Tuning table (filled by swfd_tune, e.g. from an autotuner): per kind and
// grid size, the row-segment height, a launch-shape parameter and the
// temporal depth (1: single-step kernels, 2: two-step kernels).
struct TuneEntry {
    int kind, ny, nx, seg_rows, param, depth;
};
std::vector<TuneEntry> g_tune;

const TuneEntry *find_tune(int kind, int ny, int nx)
{
    for (const TuneEntry &t : g_tune)
        if (t.kind == kind && t.ny == ny && t.nx == nx) return &t;
    return nullptr;
}

// Geometry of a two-step forward launch (work items: shot group x 64-column
// block x row segment); returns its workgroup count, gather workgroups
// included.
int pair_geometry(StepParams &p, int gather_pairs_count)
{
    static const int seg_env = std::max(2, env_int("SWFD_SEG_ROWS", kSegRows));
    static const int wpc_env = env_int("SWFD_TB_WAVES_PER_CU", 256);
    static const int spg_override = env_int("SWFD_TB_SHOTS_PER_GROUP", 0);
    const TuneEntry *t = find_tune(0, p.ny, p.nx);
    const int seg_rows = t != nullptr && t->seg_rows >= 2 ? t->seg_rows : seg_env;
    const int waves_per_cu = t != nullptr && t->param > 0 ? t->param : wpc_env;
    p.n_cb = (p.nx + kLanes - 1) / kLanes;
    p.seg = seg_rows;
    p.n_seg = std::max(1, p.ny / seg_rows);
    const long cells = (long)p.n_cb * p.n_seg;
    const long target = (long)waves_per_cu * cu_count();
    long groups = (target + cells - 1) / cells;
    if (spg_override > 0) groups = (p.n_shots + spg_override - 1) / spg_override;
    groups = std::max<long>(1, std::min<long>(groups, p.n_shots));
    p.shots_per_group = (int)((p.n_shots + groups - 1) / groups);
    p.n_groups = (p.n_shots + p.shots_per_group - 1) / p.shots_per_group;
    p.n_work = (int)(cells * p.n_groups);
    const int blocks = (p.n_work + kWavesPerBlock - 1) / kWavesPerBlock;
    p.chunk = (blocks + kXcds - 1) / kXcds;
    p.n_stencil_blocks = p.chunk * kXcds;
    const int gather_blocks = gather_pairs_count > 0 ? (gather_pairs_count + 255) / 256 : 0;
    return p.n_stencil_blocks + gather_blocks;
}

unsigned pair_flags(const StepParams &p, bool gather)
{
    unsigned f = kPoint;
    if (p.f_out2 != nullptr || p.f_out2_b != nullptr) f |= kOut2;
    if (gather) f |= kGather;
    return f;
}

template <typename K>
void launch_fwd2_variant(unsigned f, dim3 grid, hipStream_t st, K &&launch)
{
    switch (f) {
        case kPoint: launch(fwd2_kernel<kPoint>, fwd2_batch_kernel<kPoint>); break;
        case kPoint | kGather:
            launch(fwd2_kernel<kPoint | kGather>, fwd2_batch_kernel<kPoint | kGather>);
            break;
        case kPoint | kOut2:
            launch(fwd2_kernel<kPoint | kOut2>, fwd2_batch_kernel<kPoint | kOut2>);
            break;
        default:
            launch(fwd2_kernel<kPoint | kOut2 | kGather>,
                   fwd2_batch_kernel<kPoint | kOut2 | kGather>);
            break;
    }
    (void)grid;
    (void)st;
}

// Whether a row of nx columns fits one full-width workgroup.
bool wide_fits(int nx)
{
    const int off = (nx % kLanes == 1) ? 1 : 0;
    return (nx + off + kLanes - 1) / kLanes <= kMaxWideWaves;
}

// Steps per launch and form of the forward runs of a grid. A tuned entry of
// kind 2 gives the steps per launch (depth 1..4) and the form (param 1 =
// full width; more than two steps need full width); otherwise full width
// with SWFD_FWD_LEVELS (default 4) steps where a row fits, else two-step
// 64-column blocks. No fourth work buffer or SWFD_TWO_STEP=0: single steps.
struct ForwardPlan {
    int levels;
    bool wide;
};

ForwardPlan forward_plan(int ny, int nx, const float *work3)
{
    static const int env_levels = std::min(std::max(env_int("SWFD_FWD_LEVELS", 4), 1), kMaxLevels);
    static const bool env_two = env_int("SWFD_TWO_STEP", 1) != 0;
    static const bool env_wide = env_int("SWFD_WIDE", 1) != 0;
    ForwardPlan f{1, false};
    if (work3 == nullptr || !env_two) return f;
    const bool fits = wide_fits(nx);
    const TuneEntry *t = find_tune(2, ny, nx);
    if (t != nullptr && t->depth > 0) {
        f.levels = std::min(t->depth, kMaxLevels);
        f.wide = fits && f.levels >= 2 && (t->param == 1 || f.levels > 2);
        if (f.levels > 2 && !fits) f.levels = 2;
        return f;
    }
    f.wide = fits && env_wide && env_levels >= 2;
    f.levels = f.wide ? env_levels : std::min(env_levels, 2);
    return f;
}

// Geometry of a full-width forward launch; returns its workgroup count
// (gather workgroups included) and the threads per workgroup.
int wide_geometry(StepParams &p, int gather_pairs_count, int &threads)
{
    static const int seg_env = std::max(2, env_int("SWFD_WIDE_SEG_ROWS", 64));
    static const int spg_env = std::max(1, env_int("SWFD_WIDE_SHOTS_PER_GROUP", 1));
    const TuneEntry *t = find_tune(0, p.ny, p.nx);
    const bool fixed = (t != nullptr && t->seg_rows >= 2) || env_int("SWFD_WIDE_SEG_ROWS", 0) >= 2;
    int seg_rows = t != nullptr && t->seg_rows >= 2 ? t->seg_rows : seg_env;
    p.wide_off = (p.nx % kLanes == 1) ? 1 : 0;
    p.wide_waves = (p.nx + p.wide_off + kLanes - 1) / kLanes;
    threads = p.wide_waves * kLanes;
    p.shots_per_group = std::min(spg_env, p.n_shots);
    p.n_groups = (p.n_shots + p.shots_per_group - 1) / p.shots_per_group;
    // narrow rows: shorter segments until the workgroups fill half the CUs'
    // capacity (64 VGPRs: 32 wavefronts per CU)
    if (!fixed) {
        const long target = (long)cu_count() * std::max(1, 32 / p.wide_waves) / 2;
        while (seg_rows > 16 && (long)std::max(1, p.ny / seg_rows) * p.n_groups < target)
            seg_rows /= 2;
    }
    p.seg = seg_rows;
    p.n_seg = std::max(1, p.ny / seg_rows);
    p.n_work = p.n_seg * p.n_groups;
    p.chunk = (p.n_work + kXcds - 1) / kXcds;
    p.n_stencil_blocks = p.chunk * kXcds;
    const int gather_blocks =
        gather_pairs_count > 0 ? (gather_pairs_count + threads - 1) / threads : 0;
    return p.n_stencil_blocks + gather_blocks;
}

// Kernel variants: store pattern x (point source, checkpoint copies, fused
// gather). Copies take the per-level checks; gathers come with two levels
// only (receiver samples are read from the stored fields). `fn` receives the
// single-problem and the batched kernel of the variant.
template <int KL, typename Fn>
int fwdm_variant(int sm, unsigned f, Fn &&fn)
{
#define SWFD_FWDM(SMV, FV) fn(fwdm_kernel<KL, SMV, FV>, fwdm_batch_kernel<KL, SMV, FV>)
    if ((f & kBnd) != 0) {
        // boundary cells kept by footprint stores (point sources, optional
        // checkpoint copies, gathers after the launch)
        if constexpr (KL > 2) {
            if (sm == kStoreFootprint && (f & ~kOut2) == (kPoint | kBnd)) {
                if ((f & kOut2) != 0) SWFD_FWDM(kStoreFootprint, kPoint | kOut2 | kBnd);
                else SWFD_FWDM(kStoreFootprint, kPoint | kBnd);
                return SWFD_OK;
            }
        }
        return fail(SWFD_EINVAL, "boundary cells: footprint stores of 3-4 levels only");
    }
    if ((f & kLap) != 0) {
        // Laplacian histories: the fields of the last two levels are stored
        if (f != (kPoint | kLap) || sm != kStoreLastTwo)
            return fail(SWFD_EINVAL, "Laplacian history: no copies, gathers or field history");
        SWFD_FWDM(kStoreLastTwo, kPoint | kLap);
        return SWFD_OK;
    }
    if constexpr (KL == 1) {
        return fail(SWFD_EINVAL, "multi-step launch: one level only with Laplacian histories");
    } else if ((f & kOut2) != 0) {
        if ((f & kGather) != 0) {
            if constexpr (KL == 2) SWFD_FWDM(kStoreAny, kPoint | kOut2 | kGather);
            else return fail(SWFD_EINVAL, "multi-step launch: gathers need two levels");
        } else if (sm == kStoreFootprint) {
            // checkpoint copies go to their own outputs in full either way
            if constexpr (KL > 2) SWFD_FWDM(kStoreFootprint, kPoint | kOut2);
            else SWFD_FWDM(kStoreAny, kPoint | kOut2);
        } else {
            SWFD_FWDM(kStoreAny, kPoint | kOut2);
        }
        return SWFD_OK;
    }
    if ((f & kGather) != 0) {
        if constexpr (KL == 2) {
            switch (sm) {
                case kStoreAll: SWFD_FWDM(kStoreAll, kPoint | kGather); break;
                case kStoreLastTwo: SWFD_FWDM(kStoreLastTwo, kPoint | kGather); break;
                default: SWFD_FWDM(kStoreAny, kPoint | kGather); break;
            }
            return SWFD_OK;
        } else {
            return fail(SWFD_EINVAL, "multi-step launch: gathers need two levels");
        }
    }
    if constexpr (KL >= 2) {
        switch (sm) {
            case kStoreAll: SWFD_FWDM(kStoreAll, kPoint); break;
            case kStoreLastTwo: SWFD_FWDM(kStoreLastTwo, kPoint); break;
            case kStoreFootprint:
                // two levels are both stored everywhere
                if constexpr (KL > 2) SWFD_FWDM(kStoreFootprint, kPoint);
                else SWFD_FWDM(kStoreAll, kPoint);
                break;
            default: SWFD_FWDM(kStoreAny, kPoint); break;
        }
    }
#undef SWFD_FWDM
    return SWFD_OK;
}

template <typename Fn>
int fwdm_levels_variant(int levels, int sm, unsigned f, Fn &&fn)
{
    switch (levels) {
        case 1: return fwdm_variant<1>(sm, f, fn);
        case 2: return fwdm_variant<2>(sm, f, fn);
        case 3: return fwdm_variant<3>(sm, f, fn);
        default: return fwdm_variant<4>(sm, f, fn);
    }
}

// Shape of a full-width multi-step forward launch (p.levels levels; level
// j's outputs p.lv_out[j] / p.lv_out2[j], either may be NULL, the last two
// levels stored by the caller's plan): fills the geometry into p and returns
// the workgroup count, the threads per workgroup, the flags and the store
// pattern.
// Whether footprint-store forward launches may use the lean marches
// (SWFD_FWD_LEAN=0 turns them off; read at every launch).
bool fwdm_lean_enabled()
{
    const char *e = std::getenv("SWFD_FWD_LEAN");
    return e == nullptr || std::strcmp(e, "0") != 0;
}

int multi_shape(StepParams &p, int gather_pairs_count, int &threads, unsigned &f, int &sm)
{
    const int min_levels = p.hist_lap ? 1 : 2;
    if (p.src_rows == nullptr || p.levels < min_levels || p.levels > kMaxLevels ||
        !wide_fits(p.nx))
        return -fail(SWFD_EINVAL, "multi-step launch: needs point sources, 2-4 levels, nx <= 1023");
    if ((size_t)p.ny * p.nx * sizeof(float) >= (size_t)1 << 31)
        return -fail(SWFD_EINVAL, "multi-step launch: a field plane must be under 2 GiB");
    const int blocks = wide_geometry(p, gather_pairs_count, threads);
    f = kPoint;
    bool all = true, last_two = true;
    for (int j = 0; j < p.levels; ++j) {
        if (p.lv_out2[j] != nullptr) f |= kOut2;
        const bool st = p.lv_out[j] != nullptr;
        all = all && st;
        last_two = last_two && (st == (j >= p.levels - 2));
    }
    if (gather_pairs_count > 0) f |= kGather;
    if (p.hist_lap) f |= kLap;

    const bool footprint = all && p.fp_rows != nullptr && p.levels > 2;
    sm = (p.hist_lap && last_two) ? kStoreLastTwo
         : (footprint ? kStoreFootprint
                      : (all ? kStoreAll : (last_two ? kStoreLastTwo : kStoreAny)));
    // a ring is taken from the stored fields: levels kept only where receivers
    // sit keep their boundary cells too
    if (p.bnd != nullptr && sm == kStoreFootprint) f |= kBnd;
    // lean marches for footprint-store launches (SWFD_FWD_LEAN=0: the general
    // march everywhere, with the same bits)
    p.fwdm_lean = (sm == kStoreFootprint && p.wide_waves >= 2 && p.ny / p.n_seg >= 2 &&
                   fwdm_lean_enabled()) ? 1 : 0;
    return blocks;
}

// p.levels (1..4) forward steps in one full-width launch of one problem.
int launch_multi(StepParams &p, hipStream_t stream, int gather_pairs_count)
{
    int threads = 0, sm = 0;
    unsigned f = 0;
    const int blocks = multi_shape(p, gather_pairs_count, threads, f, sm);
    if (blocks < 0) return -blocks;
    const dim3 grid(blocks), block(threads);
    const int rc = fwdm_levels_variant(p.levels, sm, f, [&](auto single, auto) {
        hipLaunchKernelGGL(single, grid, block, 0, stream, p);
    });
    if (rc) return rc;
    SWFD_HIP_CHECK(hipGetLastError());
    return SWFD_OK;
}

// Two forward steps in one 64-column-block launch: outputs p.f_out (first
// step) and p.f_out_b (second); p.f_out2 / p.f_out2_b optional copies.
int launch_pair(StepParams &p, hipStream_t stream, int gather_pairs_count)
{
    if (p.src_rows == nullptr || p.f_out == nullptr || p.f_out_b == nullptr)
        return fail(SWFD_EINVAL, "two-step launch needs point sources and two outputs");
    const dim3 grid(pair_geometry(p, gather_pairs_count));
    const dim3 block(kLanes * kWavesPerBlock);
    const unsigned f = pair_flags(p, gather_pairs_count > 0);
    launch_fwd2_variant(f, grid, stream, [&](auto single, auto) {
        hipLaunchKernelGGL(single, grid, block, 0, stream, p);
    });
    SWFD_HIP_CHECK(hipGetLastError());
    return SWFD_OK;
}

// Shape of a two-step adjoint launch: workgroups of four wavefronts take one
// shot group at one (column block, row segment); each wavefront marches
// shots_per_wave shots. With imaging there are at most kMaxImageGroups shot
// groups (one partial image each).
struct Adj2Shape {
    int n_cb, n_seg, seg, n_groups, shots_per_group, shots_per_wave, acc_rows;
};

Adj2Shape adj2_shape(int ny, int nx, int n_shots)
{
    static const int seg_env = std::max(2, env_int("SWFD_ADJ_SEG_ROWS", 32));
    static const int groups_env = env_int("SWFD_ADJ_GROUPS", kMaxImageGroups);
    const TuneEntry *t = find_tune(1, ny, nx);
    int seg_rows = t != nullptr && t->seg_rows >= 2 ? t->seg_rows : seg_env;
    const int groups_param = t != nullptr && t->param > 0 ? t->param : groups_env;
    // the imaging accumulator (4 wavefronts x rows x 64 floats) must fit 64 KiB
    seg_rows = std::min(seg_rows, 32);
    Adj2Shape a{};
    a.n_cb = (nx + kLanes - 1) / kLanes;
    a.seg = seg_rows;
    a.n_seg = std::max(1, ny / seg_rows);
    a.acc_rows = (ny + a.n_seg - 1) / a.n_seg;
    const int groups = std::max(1, std::min(std::min(groups_param, kMaxImageGroups), n_shots));
    a.shots_per_group = (n_shots + groups - 1) / groups;
    a.shots_per_wave = (a.shots_per_group + kWavesPerBlock - 1) / kWavesPerBlock;
    a.shots_per_group = a.shots_per_wave * kWavesPerBlock;
    a.n_groups = (n_shots + a.shots_per_group - 1) / a.shots_per_group;
    return a;
}

// Geometry of a two-step adjoint launch; returns its workgroup count and the
// LDS bytes it needs (0 without imaging).
int adj_pair_geometry(StepParams &p, size_t &lds)
{
    const Adj2Shape a = adj2_shape(p.ny, p.nx, p.n_shots);
    p.n_cb = a.n_cb;
    p.n_seg = a.n_seg;
    p.seg = a.seg;
    p.n_groups = a.n_groups;
    p.shots_per_group = a.shots_per_group;
    p.shots_per_wave = a.shots_per_wave;
    p.acc_rows = a.acc_rows;
    p.n_work = a.n_cb * a.n_seg * a.n_groups;
    p.chunk = (p.n_work + kXcds - 1) / kXcds;
    const bool image = p.img != nullptr || p.img_b != nullptr;
    lds = image ? (size_t)kWavesPerBlock * a.acc_rows * kLanes * sizeof(float) : 0;
    return p.chunk * kXcds;
}

// Two adjoint steps in one launch: p.f_out (first step) and p.f_out_b
// (second), residuals p.resid / p.resid_b, imaging fields p.img / p.img_b
// (either may be NULL) into p.g_part.
int launch_adj_pair(StepParams &p, hipStream_t stream)
{
    if (p.resid == nullptr || p.resid_b == nullptr || p.f_out == nullptr || p.f_out_b == nullptr)
        return fail(SWFD_EINVAL, "two-step adjoint launch needs residuals and two outputs");
    const bool image = p.img != nullptr || p.img_b != nullptr;
    if (image && p.g_part == nullptr)
        return fail(SWFD_EINVAL, "two-step adjoint imaging needs the partial images");
    size_t lds = 0;
    const dim3 grid((unsigned)adj_pair_geometry(p, lds));
    const dim3 block(kLanes * kWavesPerBlock);
    if (lds > 64 * 1024) return fail(SWFD_EINVAL, "two-step adjoint: row segments too long");
    if (image)
        hipLaunchKernelGGL((adj2_kernel<kResid | kImage>), grid, block, lds, stream, p);
    else
        hipLaunchKernelGGL((adj2_kernel<kResid>), grid, block, 0, stream, p);
    SWFD_HIP_CHECK(hipGetLastError());
    return SWFD_OK;
}

// Largest LDS image of a row segment a full-width adjoint workgroup may use
// (one such workgroup per CU).
constexpr size_t kMaxAdjwLds = 128 * 1024;

// Shape of a full-width multi-step adjoint launch: rows per segment and shot
// groups (one partial image each). SWFD_ADJW_SEG_ROWS / SWFD_ADJW_GROUPS or a
// kind-1 tuning entry fix either.
//
// Recomputing kernel (adjm_kernel): about 120 VGPRs (four wavefronts per
// SIMD), so a CU holds 16 / wavefronts-per-row workgroups, fewer if their LDS
// images of the segment do not fit. A workgroup marches its segment (plus the
// 3 rows shared with each neighbour) once per shot of its group, so the
// launch takes about ceil(workgroups / (CUs x resident)) x (rows + 3) x
// shots-per-group; the shape minimises that over 8-32 row segments and 8-32
// groups, preferring longer segments and, on ties, more groups.
//
// Reconstructing kernel (adjr_kernel): a march covers the segment plus 2 KL -
// 1 = 7 rows. Measured on 256^2 - 1024^2 grids (128 and 256 shots, every
// combination of 8 / 16 / 32 rows and 8 - 128 groups), the fastest launch
// always had 32-row segments and about 3 - 4 workgroups per CU (256^2: 128
// groups of 1 shot; 384^2: 64 groups of 2; 512^2: 64 of 2; 768^2: 32 of 4;
// 1024^2 x 256 shots: 32 of 8), with whole rounds of workgroups over the CUs
// (768^2 with 43 groups, 4.03 per CU: 18 % slower than 32 groups). More
// groups per segment share its c^2 and scatter-map rows in an XCD's L2; too
// many pay a partial image each. The shape minimises |log(ceil(m) / 3.5)| +
// (ceil(m) - m) / ceil(m) + log((rows + 7) / rows), m = workgroups per CU,
// over 8-32 row segments and 8 - max_wide_groups groups (on ties, longer
// segments, then more groups).
constexpr int kMaxWideGroups = 512;
// Floats of partial images a launch may use: 32 images at 1024^2.
constexpr size_t kWideImageFloats = (size_t)32 << 20;

// Shot groups (partial images) a full-width adjoint launch on ny x nx may
// use: 32 at 1024^2 and above, up to kMaxWideGroups on small grids (512 at
// 256^2), so that small grids can fill the CUs with short shot loops.
int max_wide_groups(int ny, int nx)
{
    const size_t plane = (size_t)ny * nx;
    return (int)std::max<size_t>(32, std::min<size_t>(kMaxWideGroups, kWideImageFloats / plane));
}

struct AdjwShape {
    int n_seg, seg, n_groups, shots_per_group, acc_rows;
};

// recon: the shape of the reconstructing kernel (adjr_kernel), else of the
// recomputing one (adjm_kernel).
AdjwShape adjw_shape(int ny, int nx, int n_shots, bool recon)
{
    static const int seg_env = env_int("SWFD_ADJW_SEG_ROWS", 0);
    static const int groups_env = env_int("SWFD_ADJW_GROUPS", 0);
    const TuneEntry *t = find_tune(1, ny, nx);
    const int seg_fixed = t != nullptr && t->seg_rows >= 2 ? t->seg_rows : seg_env;
    const int groups_fixed = t != nullptr && t->param > 0 ? t->param : groups_env;
    const int max_groups =
        std::max(1, std::min(recon ? max_wide_groups(ny, nx) : 32, n_shots));
    const int off = (nx % kLanes == 1) ? 1 : 0;
    const int waves = (nx + off + kLanes - 1) / kLanes;
    const int max_rows = std::max(1, (int)(kMaxAdjwLds / (sizeof(float) * (size_t)nx)));
    auto n_seg_of = [&](int seg) {
        const int n = std::max(1, ny / seg);
        return std::max(n, (ny + max_rows - 1) / max_rows);
    };
    auto cost = [&](int seg, int groups) -> double {
        const int n_seg = n_seg_of(seg);
        const int rows = (ny + n_seg - 1) / n_seg;
        const long spg = (n_shots + groups - 1) / groups;
        if (recon) {
            const long used = (n_shots + spg - 1) / spg;  // groups actually formed
            const double m = (double)n_seg * used / cu_count();
            const double cm = std::ceil(m);
            return std::fabs(std::log(cm / 3.5)) + (cm - m) / cm +
                   std::log((rows + 2.0 * kMaxLevels - 1.0) / rows);
        }
        const size_t lds = (size_t)rows * nx * sizeof(float) + 4096;
        const long resident = std::max(1, std::min(16 / waves, (int)((160 * 1024) / lds)));
        const long capacity = (long)cu_count() * resident;
        const long rounds = ((long)n_seg * groups + capacity - 1) / capacity;
        return (double)(rounds * (rows + 3) * spg);
    };
    int best_seg = seg_fixed >= 2 ? seg_fixed : 32;
    int best_groups = groups_fixed > 0 ? std::min(groups_fixed, max_groups) : std::min(8, max_groups);
    double best = cost(best_seg, best_groups);
    static const int kGroupChoices[7] = {8, 16, 32, 64, 128, 256, kMaxWideGroups};
    for (int seg = 32; seg >= 8; seg /= 2) {
        if (seg_fixed >= 2 && seg != seg_fixed) continue;
        for (int gi = 0; gi < 7; ++gi) {
            if (!recon && kGroupChoices[gi] > 32) break;
            const int groups = groups_fixed > 0 ? std::min(groups_fixed, max_groups)
                                                : std::min(kGroupChoices[gi], max_groups);
            const double c = cost(seg, groups);
            // on a tie, more groups: shorter shot loops per workgroup, and the
            // rounds overlap one workgroup's start with another's end (1024^2,
            // 128 shots: 32 rows x 16 groups 1.35 ms per launch, x 8 groups 1.40)
            if (c < best - 1e-12 ||
                (c <= best + 1e-12 && seg == best_seg && groups > best_groups)) {
                best = c;
                best_seg = seg;
                best_groups = groups;
            }
        }
    }
    AdjwShape a{};
    a.seg = best_seg;
    a.n_seg = n_seg_of(best_seg);
    a.acc_rows = (ny + a.n_seg - 1) / a.n_seg;
    a.shots_per_group = (n_shots + best_groups - 1) / best_groups;
    a.n_groups = (n_shots + a.shots_per_group - 1) / a.shots_per_group;
    return a;
}

// Imaging variant of a full-width adjoint launch: 0 none, 1 fields, 2
// Laplacians (history of Laplacians).
int adjm_image(const StepParams &p)
{
    bool image = false, every = true;
    for (int j = 0; j < p.levels; ++j) {
        image = image || p.lv_in[j] != nullptr;
        every = every && p.lv_in[j] != nullptr;
    }
    return !image ? 0 : (p.hist_lap ? 2 : (every ? 3 : 1));
}

// Shape of a full-width adjoint launch of p.levels (2..4) steps: outputs of
// the last two levels in p.lv_out[levels - 2], p.lv_out[levels - 1];
// residuals p.lv_resid[j]; imaging fields p.lv_in[j] (NULL: level not
// imaged). Fills the geometry into p; returns the workgroup count (negative
// error code on failure), the threads per workgroup and the LDS bytes.
int adjm_shape(StepParams &p, int &threads, size_t &lds)
{
    const int L = p.levels;
    if (L < 2 || L > kMaxLevels || !wide_fits(p.nx) || p.cell_map == nullptr)
        return -fail(SWFD_EINVAL, "multi-step adjoint: 2-4 levels, nx <= 1023, scatter table");
    if ((size_t)p.ny * p.nx * sizeof(float) >= (size_t)1 << 31)
        return -fail(SWFD_EINVAL, "multi-step adjoint: a field plane must be under 2 GiB");
    for (int j = 0; j < L; ++j)
        if (p.lv_resid[j] == nullptr) return -fail(SWFD_EINVAL, "multi-step adjoint: residual");
    if (p.lv_out[L - 2] == nullptr || p.lv_out[L - 1] == nullptr)
        return -fail(SWFD_EINVAL, "multi-step adjoint needs the last two outputs");
    const bool image = adjm_image(p) != 0;
    if (image && p.g_part == nullptr)
        return -fail(SWFD_EINVAL, "multi-step adjoint imaging needs the partial images");
    const AdjwShape a = adjw_shape(p.ny, p.nx, p.n_shots, false);
    p.wide_off = (p.nx % kLanes == 1) ? 1 : 0;
    p.wide_waves = (p.nx + p.wide_off + kLanes - 1) / kLanes;
    p.seg = a.seg;
    p.n_seg = a.n_seg;
    p.n_groups = a.n_groups;
    p.shots_per_group = a.shots_per_group;
    p.acc_rows = a.acc_rows;
    p.n_work = p.n_seg * p.n_groups;
    p.chunk = (p.n_work + kXcds - 1) / kXcds;
    p.n_stencil_blocks = p.chunk * kXcds;
    lds = image ? (size_t)a.acc_rows * p.nx * sizeof(float) : 0;
    if (lds > kMaxAdjwLds)
        return -fail(SWFD_EINVAL, "multi-step adjoint: row segments too long for the LDS image");
    threads = p.wide_waves * kLanes;
    return p.n_stencil_blocks;
}

// The kernels of a full-width adjoint variant, handed to fn(single,
// batched), after allowing them dynamic LDS beyond 64 KiB (once each).
template <int KL, int IMG, typename Fn>
int adjm_variant_img(Fn &&fn)
{
    if constexpr (IMG != 0) {
        static bool big = false;
        if (!big) {
            SWFD_HIP_CHECK(hipFuncSetAttribute(
                reinterpret_cast<const void *>(&adjm_kernel<KL, IMG>),
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxAdjwLds));
            SWFD_HIP_CHECK(hipFuncSetAttribute(
                reinterpret_cast<const void *>(&adjm_batch_kernel<KL, IMG>),
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxAdjwLds));
            big = true;
        }
    }
    fn(adjm_kernel<KL, IMG>, adjm_batch_kernel<KL, IMG>);
    return SWFD_OK;
}

template <int KL, typename Fn>
int adjm_variant(int img, Fn &&fn)
{
    switch (img) {
        case 0: return adjm_variant_img<KL, 0>(fn);
        case 1: return adjm_variant_img<KL, 1>(fn);
        case 3: return adjm_variant_img<KL, 3>(fn);
        default: return adjm_variant_img<KL, 2>(fn);
    }
}

template <typename Fn>
int adjm_levels_variant(int levels, int img, Fn &&fn)
{
    switch (levels) {
        case 2: return adjm_variant<2>(img, fn);
        case 3: return adjm_variant<3>(img, fn);
        default: return adjm_variant<4>(img, fn);
    }
}

// p.levels (2..4) adjoint steps in one full-width launch of one problem.
int launch_adj_multi(StepParams &p, hipStream_t stream)
{
    int threads = 0;
    size_t lds = 0;
    const int blocks = adjm_shape(p, threads, lds);
    if (blocks < 0) return -blocks;
    const dim3 grid((unsigned)blocks), block((unsigned)threads);
    const int rc = adjm_levels_variant(p.levels, adjm_image(p), [&](auto single, auto) {
        hipLaunchKernelGGL(single, grid, block, lds, stream, p);
    });
    if (rc) return rc;
    SWFD_HIP_CHECK(hipGetLastError());
    return SWFD_OK;
}

// Dynamic LDS a reconstructing adjoint workgroup may use: the segment image
// and the staged boundary columns (one such workgroup per CU).
constexpr size_t kMaxAdjrLds = 156 * 1024;

// Shape of a reconstructing adjoint launch of p.levels (1..4) steps (the
// geometry of the full-width adjoint); returns the workgroup count (negative
// error code on failure), the threads per workgroup and the LDS bytes.
int adjr_shape(StepParams &p, int &threads, size_t &lds)
{
    const int L = p.levels;
    if (L < 1 || L > kMaxLevels || !wide_fits(p.nx) || p.cell_map == nullptr)
        return -fail(SWFD_EINVAL, "reconstructing adjoint: 1-4 levels, nx <= 1023, scatter table");
    if ((size_t)p.ny * p.nx * sizeof(float) >= (size_t)1 << 31)
        return -fail(SWFD_EINVAL, "reconstructing adjoint: a field plane must be under 2 GiB");
    for (int j = 0; j < L; ++j)
        if (p.lv_resid[j] == nullptr) return -fail(SWFD_EINVAL, "reconstructing adjoint: residual");
    const int lo = std::max(L - 2, 0);
    for (int j = lo; j < L; ++j)
        if (p.lv_out[j] == nullptr || p.lv_rout[j] == nullptr)
            return -fail(SWFD_EINVAL, "reconstructing adjoint needs the last outputs");
    if (p.g_part == nullptr || p.r_next == nullptr || p.r_cur == nullptr || p.bnd == nullptr ||
        p.src_rows == nullptr || p.src_cols == nullptr || p.src_values == nullptr)
        return -fail(SWFD_EINVAL, "reconstructing adjoint: fields, ring, sources or images missing");
    const AdjwShape a = adjw_shape(p.ny, p.nx, p.n_shots, true);
    p.wide_off = (p.nx % kLanes == 1) ? 1 : 0;
    p.wide_waves = (p.nx + p.wide_off + kLanes - 1) / kLanes;
    p.seg = a.seg;
    p.n_seg = a.n_seg;
    p.n_groups = a.n_groups;
    p.shots_per_group = a.shots_per_group;
    p.acc_rows = a.acc_rows;
    p.bcol_rows = a.acc_rows + 3 * L;
    p.n_work = p.n_seg * p.n_groups;
    p.chunk = (p.n_work + kXcds - 1) / kXcds;
    p.n_stencil_blocks = p.chunk * kXcds;
    // image, staged boundary columns, staged rows 0 / H - 1 (both only when a
    // segment's march can reach both)
    const size_t brows = (p.ny <= a.acc_rows + 3 * L ? 2 : 1) * (size_t)L * p.nx;
    lds = ((size_t)a.acc_rows * p.nx + (size_t)L * 2 * p.bcol_rows + brows) * sizeof(float);
    if (lds > kMaxAdjrLds)
        return -fail(SWFD_EINVAL, "reconstructing adjoint: row segments too long for the LDS");
    threads = p.wide_waves * kLanes;
    return p.n_stencil_blocks;
}

// fn(single, batched) with the reconstructing adjoint kernels of KL steps
// (after raising their dynamic-LDS limit once).
template <int KL, bool EVERY, typename Fn>
int adjr_variant_kl(Fn &&fn)
{
    static bool big = false;
    if (!big) {
        SWFD_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&adjr_kernel<KL, EVERY>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)kMaxAdjrLds));
        SWFD_HIP_CHECK(hipFuncSetAttribute(
            reinterpret_cast<const void *>(&adjr_batch_kernel<KL, EVERY>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxAdjrLds));
        big = true;
    }
    fn(adjr_kernel<KL, EVERY>, adjr_batch_kernel<KL, EVERY>);
    return SWFD_OK;
}

template <typename Fn>
int adjr_variant(int levels, bool every, Fn &&fn)
{
    switch (levels * 2 + (every ? 1 : 0)) {
        case 2: return adjr_variant_kl<1, false>(fn);
        case 3: return adjr_variant_kl<1, true>(fn);
        case 4: return adjr_variant_kl<2, false>(fn);
        case 5: return adjr_variant_kl<2, true>(fn);
        case 6: return adjr_variant_kl<3, false>(fn);
        case 7: return adjr_variant_kl<3, true>(fn);
        case 8: return adjr_variant_kl<4, false>(fn);
        default: return adjr_variant_kl<4, true>(fn);
    }
}

// Whether reconstructing launches may use the lean marches
// (SWFD_ADJR_LEAN=0 turns them off; read at every launch).
bool adjr_lean_enabled()
{
    const char *e = std::getenv("SWFD_ADJR_LEAN");
    return e == nullptr || std::strcmp(e, "0") != 0;
}

// Whether every level of a reconstructing launch images (its steps are all at
// least p.img_from), and the lean-march flag (also: every rebuilt field of
// step >= 0, segments of at least two rows and rows of at least two
// wavefronts; SWFD_ADJR_LEAN=0 sends every wavefront through the general
// march, with the same bits). Call after adjr_shape.
bool adjr_every(const StepParams &p) { return p.src_index - p.levels + 1 >= p.img_from; }

void adjr_set_lean(StepParams &p)
{
    p.adjr_lean = (adjr_every(p) && p.src_index - p.levels - 1 >= 0 && p.ny / p.n_seg >= 2 &&
                   p.wide_waves >= 2 && adjr_lean_enabled()) ? 1 : 0;
}

// p.levels (1..4) adjoint steps with the forward field rebuilt in one
// full-width launch; every level whose step is at least p.img_from images.
int launch_adj_recon(StepParams &p, hipStream_t st)
{
    int threads = 0;
    size_t lds = 0;
    const int blocks = adjr_shape(p, threads, lds);
    if (blocks < 0) return -blocks;
    const dim3 grid((unsigned)blocks), block((unsigned)threads);
    adjr_set_lean(p);
    const int rc = adjr_variant(p.levels, adjr_every(p), [&](auto single, auto) {
        hipLaunchKernelGGL(single, grid, block, lds, st, p);
    });
    if (rc) return rc;
    SWFD_HIP_CHECK(hipGetLastError());
    return SWFD_OK;
}

// Number of imaging partial slices a step of this shape uses.
int image_groups(int ny, int nx, int n_shots)
{
    return choose_shape(ny, nx, n_shots, true).n_groups;
}

// Imaging of the field `lam` (S, ny, nx) against `img` into g_out.
int launch_image(const swfd_grid *grid, int n_shots, const float *lam, const float *img,
                 const float *g_in, float *g_out, float scale, hipStream_t stream,
                 bool lap = false)
{
    StepParams q;
    base_params(q, grid, nullptr, n_shots);
    q.hist_lap = lap ? 1 : 0;
    q.f_cur = lam;
    q.img = img;
    q.g_in = g_in;
    q.g_out = g_out;
    q.img_scale = scale;
    const size_t plane = (size_t)grid->ny * grid->nx;
    hipLaunchKernelGGL(image_kernel, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0,
                       stream, q);
    SWFD_HIP_CHECK(hipGetLastError());
    return SWFD_OK;
}

int check_grid(const swfd_grid *g)
{
    if (g == nullptr) return fail(SWFD_EINVAL, "grid is NULL");
    if (g->ny < 3 || g->nx < 3)
        return fail(SWFD_EINVAL, "grid must be at least 3 x 3 (got " + std::to_string(g->ny) +
                                     " x " + std::to_string(g->nx) + ")");
    return SWFD_OK;
}

int check_medium(const swfd_medium *m)
{
    if (m == nullptr || m->c2 == nullptr || m->c2dt2 == nullptr || m->mur == nullptr)
        return fail(SWFD_EINVAL, "medium coefficients are missing");
    return SWFD_OK;
}

int check_receivers(const swfd_receivers *r, bool need_table)
{
    if (r == nullptr || r->n < 1 || r->row0 == nullptr || r->col0 == nullptr ||
        r->weights == nullptr)
        return fail(SWFD_EINVAL, "receivers are missing");
    if (need_table && (r->cell_map == nullptr || r->entry_start == nullptr ||
                       r->pair_rec == nullptr || r->pair_w == nullptr ||
                       r->entry_cell == nullptr))
        return fail(SWFD_EINVAL, "receiver scatter table is missing");
    return SWFD_OK;
}

hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

int copy_field(float *dst, const float *src, size_t n, hipStream_t stream)
{
    if (dst == src) return SWFD_OK;
    SWFD_HIP_CHECK(
        hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
    return SWFD_OK;
}

// Move the state (p_old, p_cur) into (dst_old, dst_cur) using `spare` as a
// temporary when the two destinations hold each other's data.
int settle_state(const float *p_old, const float *p_cur, float *dst_old, float *dst_cur,
                 float *const work[4], size_t n, hipStream_t stream)
{
    int rc;
    if (dst_old == nullptr && dst_cur == nullptr) return SWFD_OK;
    if (dst_old == p_old && dst_cur == p_cur) return SWFD_OK;
    if (dst_old != nullptr && dst_cur != nullptr && dst_old == p_cur && dst_cur == p_old) {
        float *spare = nullptr;
        for (int i = 0; i < 3; ++i)
            if (work[i] != dst_old && work[i] != dst_cur) spare = work[i];
        if ((rc = copy_field(spare, p_cur, n, stream))) return rc;
        if ((rc = copy_field(dst_old, p_old, n, stream))) return rc;
        return copy_field(dst_cur, spare, n, stream);
    }
    if (dst_old != nullptr && dst_old == p_cur) {
        // dst_old currently holds the newest state: move it first
        if (dst_cur != nullptr && (rc = copy_field(dst_cur, p_cur, n, stream))) return rc;
        return copy_field(dst_old, p_old, n, stream);
    }
    if (dst_old != nullptr && (rc = copy_field(dst_old, p_old, n, stream))) return rc;
    if (dst_cur != nullptr && (rc = copy_field(dst_cur, p_cur, n, stream))) return rc;
    return SWFD_OK;
}

// A work buffer that is none of a, b, c (work[3] may be NULL).
float *free_work(float *const work[4], const float *a, const float *b,
                 const float *c = nullptr)
{
    for (int i = 0; i < 4; ++i)
        if (work[i] != nullptr && work[i] != a && work[i] != b && work[i] != c) return work[i];
    return nullptr;
}


// ---------------------------------------------------------------------------
// Runs as state machines: one run = the launches of one swfd_forward or
// swfd_adjoint call. swfd_forward / swfd_adjoint drive one run; the batch
// entry points drive several in lockstep, packing the two-step launches of
// all runs at each tick into one batched kernel launch.
// ---------------------------------------------------------------------------

int check_forward_args(const swfd_grid *grid, const swfd_medium *medium,
                       const swfd_receivers *rec, const swfd_forward_args *a)
{
    int rc;
    if ((rc = check_grid(grid)) || (rc = check_medium(medium))) return rc;
    if (a == nullptr) return fail(SWFD_EINVAL, "forward args are NULL");
    if (a->n_shots < 1 || a->n_steps < 0 || a->step_begin < 0)
        return fail(SWFD_EINVAL, "forward: bad n_shots / n_steps / step_begin");
    if (a->init_old == nullptr || a->init_u == nullptr || a->work0 == nullptr ||
        a->work1 == nullptr || a->work2 == nullptr)
        return fail(SWFD_EINVAL, "forward: state or work buffers are NULL");
    if (a->src_rows == nullptr || a->src_cols == nullptr || a->src_values == nullptr)
        return fail(SWFD_EINVAL, "forward: point sources are required");
    if (a->samples != nullptr && (rc = check_receivers(rec, false))) return rc;
    if (a->hist != nullptr && (a->hist_every < 1 || a->hist_slots < 1))
        return fail(SWFD_EINVAL, "forward: hist_every and hist_slots must be positive");
    if (a->ckpt != nullptr && (a->ckpt_every < 2 || a->ckpt_slots < 2))
        return fail(SWFD_EINVAL, "forward: ckpt_every must be >= 2 and ckpt_slots >= 2");
    return SWFD_OK;
}

bool two_step_allowed(int ny, int nx, const float *work3)
{
    static const bool env = env_int("SWFD_TWO_STEP", 1) != 0;
    const TuneEntry *t = find_tune(2, ny, nx);
    return env && work3 != nullptr && (t == nullptr || t->depth != 1);
}

struct LaunchGroup;

struct ForwardRun {
    const swfd_grid *grid = nullptr;
    const swfd_medium *medium = nullptr;
    const swfd_receivers *rec = nullptr;
    const swfd_forward_args *a = nullptr;
    size_t field = 0, sr = 0;
    float *work[4] = {nullptr, nullptr, nullptr, nullptr};
    int max_levels = 1;  // steps per launch
    bool wide = false;   // full-width launches (else 64-column blocks, two steps)
    bool lap = false;    // history slots receive Laplacians of the fields
    // samples of every level gathered by a kernel after each launch (full-width
    // runs with samples and the two extra work fields; else the gather of a
    // launch's outputs is fused into the next launch)
    bool sep_gather = false;
    const float *p_old = nullptr, *p_cur = nullptr;
    int k = 0;
    // fields whose receiver samples are still to be taken (the last one or
    // two outputs); the next launch, which reads them anyway, gathers them
    const float *pend_field[2] = {nullptr, nullptr};
    int pend_step[2] = {0, 0};
    int n_pend = 0;

    // batched: lockstep tables (no Laplacian histories)
    int init(const swfd_grid *g, const swfd_medium *m, const swfd_receivers *r,
             const swfd_forward_args *args, bool batched = false)
    {
        int rc;
        if ((rc = check_forward_args(g, m, r, args))) return rc;
        grid = g;
        medium = m;
        rec = r;
        a = args;
        field = (size_t)g->ny * g->nx * a->n_shots;
        sr = (size_t)a->n_shots * (a->samples != nullptr ? r->n : 0);
        work[0] = a->work0;
        work[1] = a->work1;
        work[2] = a->work2;
        work[3] = a->work3;
        const ForwardPlan plan = forward_plan(g->ny, g->nx, a->work3);
        max_levels = plan.levels;
        wide = plan.wide;
        // receiver samples are gathered from stored fields: more than two levels
        // per launch need two more fields to keep the intermediate levels (without
        // them, launches take two steps and the next launch gathers their outputs)
        sep_gather = wide && a->samples != nullptr && max_levels > 2 &&
                     a->work4 != nullptr && a->work5 != nullptr;
        if (!wide || (a->samples != nullptr && !sep_gather)) max_levels = std::min(max_levels, 2);
        lap = a->hist_lap != 0;
        if (lap && batched)
            return fail(SWFD_EINVAL, "Laplacian history: not in batched runs");
        if (a->bnd != nullptr) {
            if (lap) return fail(SWFD_EINVAL, "boundary ring: not with Laplacian histories");
            if (a->bnd_first > a->step_begin ||
                (long)a->bnd_first + a->bnd_slots < (long)a->step_begin + a->n_steps)
                return fail(SWFD_EINVAL, "boundary ring: its slots must cover the run's steps");
            // the ring is taken from each level's stored field
            if (wide && max_levels > 2 && (a->work4 == nullptr || a->work5 == nullptr))
                return fail(SWFD_EINVAL, "boundary ring needs work4 and work5");
        }
        if (lap && (!wide || max_levels < 2))
            return fail(SWFD_EINVAL, "Laplacian history needs full-width multi-step launches "
                                     "(see swfd_lap_history) and work3");
        if (lap && (a->hist == nullptr || a->samples != nullptr || a->ckpt != nullptr))
            return fail(SWFD_EINVAL, "Laplacian history: needs hist, no samples or checkpoints");
        p_old = a->init_old;
        p_cur = a->init_u;
        return SWFD_OK;
    }
    bool done() const { return k >= a->n_steps; }
    int next_levels() const { return std::min(max_levels, a->n_steps - k); }
    bool next_is_pair() const { return next_levels() >= 2; }
    bool next_is_wide() const { return wide && (next_levels() >= 2 || lap); }
    // output of step n: its history slot, else a work buffer chosen later
    float *hist_out(int n) const
    {
        if (a->hist != nullptr && n >= a->hist_first && (n - a->hist_first) % a->hist_every == 0) {
            const int slot = (n - a->hist_first) / a->hist_every;
            if (slot < a->hist_slots) return a->hist + (size_t)slot * field;
        }
        return nullptr;
    }
    // checkpoint copy of the output of step n: the state before step jK is
    // (output of jK-2, output of jK-1)
    float *ckpt_out(int n) const
    {
        if (a->ckpt == nullptr) return nullptr;
        const int K = a->ckpt_every;
        const int j2 = n + 2, j1 = n + 1;
        int slot = -1;
        if (j2 % K == 0 && j2 / K >= 1) slot = 2 * (j2 / K - 1);
        else if (j1 % K == 0 && j1 / K >= 1) slot = 2 * (j1 / K - 1) + 1;
        return (slot >= 0 && slot < a->ckpt_slots) ? a->ckpt + (size_t)slot * field : nullptr;
    }
    // Parameters of the next launch (next_levels() steps); returns the number
    // of fused gathers (pairs of shot and receiver).
    int build(StepParams &p)
    {
        const int n = a->step_begin + k;
        base_params(p, grid, medium, a->n_shots);
        attach_receivers(p, a->samples != nullptr ? rec : nullptr);
        p.f_old = p_old;
        p.f_cur = p_cur;
        p.src_rows = a->src_rows;
        p.src_cols = a->src_cols;
        p.src_values = a->src_values;
        p.src_w = a->src_weights;
        p.src_index = n;
        p.bnd = a->bnd;
        p.bnd_first = a->bnd_first;
        p.bnd_slots = a->bnd_slots;
        int gathers = 0;
        if (a->samples != nullptr && n_pend > 0 && !sep_gather) {
            p.gat_field = pend_field[0];
            p.gat_out = a->samples + (size_t)pend_step[0] * sr;
            if (n_pend == 2) {
                p.gat_field2 = pend_field[1];
                p.gat_out2 = a->samples + (size_t)pend_step[1] * sr;
            }
            gathers = n_pend * a->n_shots * rec->n;
        }
        const int L = next_levels();
        p.levels = L;
        for (int j = 0; j < kMaxLevels; ++j) p.lv_out[j] = p.lv_out2[j] = p.lv_lap[j] = nullptr;
        // the last two levels become the next state: a history slot or a
        // work buffer other than the current state
        const float *taken = nullptr;
        if (lap) {
            // level j's input is the output of step n + j - 1: its Laplacian
            // goes to that step's slot; the fields stay in work buffers
            p.hist_lap = 1;
            for (int j = 0; j < L; ++j) {
                p.lv_lap[j] = hist_out(n + j - 1);
                if (j >= L - 2) {
                    p.lv_out[j] = free_work(work, p_old, p_cur, taken);
                    taken = p.lv_out[j];
                }
            }
            p.f_out = p.lv_out[0];
            if (L >= 2) p.f_out_b = p.lv_out[1];
            if (L == 1) p.f_out = p.lv_out[0];
            return gathers;
        }
        for (int j = 0; j < L; ++j) {
            float *out = hist_out(n + j);
            if (out == nullptr && j >= L - 2) {
                out = free_work(work, p_old, p_cur, taken);
                taken = out;
            }
            // a ring is taken from the stored fields: every level needs an output
            if (out == nullptr && (sep_gather || a->bnd != nullptr))
                out = j == 0 ? a->work4 : a->work5;
            p.lv_out[j] = out;
            p.lv_out2[j] = ckpt_out(n + j);
            if (sep_gather) p.lv_samples[j] = a->samples + (size_t)(n + j) * sr;
        }
        // levels before the last two that only feed the gather (no history
        // slot) need their fields only where the receivers sit
        bool gather_only = sep_gather && rec != nullptr && rec->footprint_rows != nullptr;
        for (int j = 0; j + 2 < L; ++j) gather_only = gather_only && hist_out(n + j) == nullptr;
        if (gather_only) p.fp_rows = rec->footprint_rows;
        p.f_out = p.lv_out[0];
        p.f_out2 = p.lv_out2[0];
        if (L >= 2) {
            p.f_out_b = p.lv_out[1];
            p.f_out2_b = p.lv_out2[1];
        }
        return gathers;
    }
    // boundary ring of the steps of a launch, from its stored fields (every
    // level of a full-width launch has an output; footprint stores keep the
    // boundary cells)
    // (field, ring slot) of each level of the next launch (call before
    // advance); returns the count, or a negative error code
    int boundary_copies(const StepParams &p, const float **fields, float **slots) const
    {
        if (a->bnd == nullptr) return 0;
        const int n = a->step_begin + k;
        int L = 1;
        if (next_is_wide()) {
            L = p.levels;
            for (int j = 0; j < L; ++j) fields[j] = p.lv_out[j];
        } else {
            fields[0] = p.f_out;
            if (next_is_pair()) {
                fields[1] = p.f_out_b;
                L = 2;
            }
        }
        const size_t total = (size_t)a->n_shots * bnd_floats(grid->ny, grid->nx);
        for (int j = 0; j < L; ++j) {
            if (fields[j] == nullptr)
                return -fail(SWFD_EINVAL, "boundary ring: a level has no output");
            slots[j] = a->bnd + (size_t)(n + j - a->bnd_first) * total;
        }
        return L;
    }
    int launch_boundary(const StepParams &p, hipStream_t st) const
    {
        BndLevels lv = {};
        const int L = boundary_copies(p, lv.field, lv.slot);
        if (L <= 0) return -L;
        const size_t total = (size_t)a->n_shots * bnd_floats(grid->ny, grid->nx);
        // every level in one launch: the levels' strided column reads overlap
        hipLaunchKernelGGL(bnd_extract_kernel, dim3((unsigned)((total + 255) / 256), (unsigned)L),
                           dim3(256), 0, st, lv, a->n_shots, grid->ny, grid->nx);
        SWFD_HIP_CHECK(hipGetLastError());
        return SWFD_OK;
    }
    // the ring copies of the next launch as batched-extraction entries
    int boundary_entries(const StepParams &p, LaunchGroup &g) const;
    // the gather of a launch's levels when they are gathered after it
    int launch_gather(const StepParams &p, hipStream_t st) const
    {
        if (!sep_gather) return SWFD_OK;
        const int pairs = p.levels * a->n_shots * rec->n;
        hipLaunchKernelGGL(gather_levels_kernel, dim3((pairs + kGatherThreads - 1) / kGatherThreads),
                           dim3(kGatherThreads), 0, st, p);
        SWFD_HIP_CHECK(hipGetLastError());
        return SWFD_OK;
    }
    void advance(const StepParams &p)
    {
        const int n = a->step_begin + k;
        const int L = p.levels;
        if (sep_gather) {
            p_old = L >= 2 ? p.lv_out[L - 2] : p_cur;
            p_cur = p.lv_out[L - 1];
            n_pend = 0;
            k += L;
            return;
        }
        if (L >= 2) {
            pend_field[0] = p.lv_out[L - 2];
            pend_step[0] = n + L - 2;
            pend_field[1] = p.lv_out[L - 1];
            pend_step[1] = n + L - 1;
            n_pend = 2;
            p_old = p.lv_out[L - 2];
            p_cur = p.lv_out[L - 1];
        } else {
            pend_field[0] = p.f_out;
            pend_step[0] = n;
            n_pend = 1;
            p_old = p_cur;
            p_cur = p.f_out;
        }
        k += L;
    }
    int finish(void *stream)
    {
        int rc;
        if (a->samples != nullptr) {
            for (int i = 0; i < n_pend; ++i)
                if ((rc = swfd_gather(grid, rec, a->n_shots, pend_field[i],
                                      a->samples + (size_t)pend_step[i] * sr, stream)))
                    return rc;
        }
        return settle_state(p_old, p_cur, a->final_old, a->final_u, work, field,
                            as_stream(stream));
    }
};

int check_adjoint_args(const swfd_grid *grid, const swfd_medium *medium,
                       const swfd_receivers *rec, const swfd_adjoint_args *a)
{
    int rc;
    if ((rc = check_grid(grid)) || (rc = check_medium(medium))) return rc;
    if (a == nullptr) return fail(SWFD_EINVAL, "adjoint args are NULL");
    if (a->n_shots < 1 || a->step_lo < 0 || a->step_hi < a->step_lo)
        return fail(SWFD_EINVAL, "adjoint: bad n_shots / step range");
    if (a->init_old == nullptr || a->init_u == nullptr || a->work0 == nullptr ||
        a->work1 == nullptr || a->work2 == nullptr)
        return fail(SWFD_EINVAL, "adjoint: state or work buffers are NULL");
    if (a->residual == nullptr) return fail(SWFD_EINVAL, "adjoint: residual is NULL");
    if ((rc = check_receivers(rec, true))) return rc;
    if (a->hist != nullptr && (a->grad == nullptr || a->hist_every < 1 || a->hist_slots < 1))
        return fail(SWFD_EINVAL, "adjoint: imaging needs grad, hist_every and hist_slots");
    if (a->hist != nullptr && a->workspace == nullptr)
        return fail(SWFD_EINVAL, "adjoint: imaging needs the workspace");
    return SWFD_OK;
}

// Step m images lam' (after step m) against the forward field after step
// m-1. Two-step launches image their own outputs (steps n and n-1); a single
// step images the step before it (its input lam), and the last step is
// imaged after the run. With two-step launches only the last step (when the
// count is odd) runs alone, so no step is imaged twice.
// Steps per launch and form of the adjoint runs of a grid: a tuned kind-3
// entry (depth = steps per launch, param 1 = full width), else blocks when
// the kind-2 entry asks for two-step blocks, else full-width launches of
// SWFD_ADJ_LEVELS (default 4) steps where a row fits and the scatter table is
// there. No fourth work buffer, SWFD_TWO_STEP=0 or a kind-2 depth of 1:
// single steps.
ForwardPlan adjoint_plan(int ny, int nx, const float *work3, const swfd_receivers *rec)
{
    static const int env_levels = std::min(std::max(env_int("SWFD_ADJ_LEVELS", 4), 1), kMaxLevels);
    static const bool env_wide = env_int("SWFD_WIDE", 1) != 0;
    ForwardPlan f{1, false};
    if (!two_step_allowed(ny, nx, work3)) return f;
    const bool fits = wide_fits(nx) && rec != nullptr && rec->cell_map != nullptr;
    const TuneEntry *t3 = find_tune(3, ny, nx);
    if (t3 != nullptr && t3->depth >= 1) {
        f.levels = std::min(t3->depth, kMaxLevels);
        f.wide = fits && f.levels >= 2 && (t3->param == 1 || f.levels > 2);
        if (!f.wide) f.levels = std::min(f.levels, 2);
        return f;
    }
    const TuneEntry *t = find_tune(2, ny, nx);
    if (t != nullptr && t->depth == 2 && t->param != 1) return ForwardPlan{2, false};
    f.wide = fits && env_wide && env_levels >= 2;
    f.levels = f.wide ? env_levels : std::min(env_levels, 2);
    return f;
}

struct AdjointRun {
    const swfd_grid *grid = nullptr;
    const swfd_medium *medium = nullptr;
    const swfd_receivers *rec = nullptr;
    const swfd_adjoint_args *a = nullptr;
    size_t field = 0, plane = 0, sr = 0;
    float *work[4] = {nullptr, nullptr, nullptr, nullptr};
    int max_levels = 1;  // steps per launch
    bool wide = false;   // full-width launches (else 64-column blocks, two steps)
    bool last_multi = false;  // the last launch imaged its own outputs
    bool lap = false;         // history slots hold Laplacians of the fields
    int n_part = 0;
    bool partials_used = false;
    const float *p_old = nullptr, *p_cur = nullptr;
    int n = 0;  // next step to run (descending)
    // forward fields rebuilt in the launches: the pair (after step n, after
    // step n - 1) and the buffers the reconstructed pairs rotate through
    bool recon = false;
    float *rwork[4] = {nullptr, nullptr, nullptr, nullptr};
    const float *r_next = nullptr, *r_cur = nullptr;

    // batched: lockstep tables (no Laplacian histories)
    int init(const swfd_grid *g, const swfd_medium *m, const swfd_receivers *r,
             const swfd_adjoint_args *args, bool batched = false)
    {
        int rc;
        if ((rc = check_adjoint_args(g, m, r, args))) return rc;
        grid = g;
        medium = m;
        rec = r;
        a = args;
        field = (size_t)g->ny * g->nx * a->n_shots;
        plane = (size_t)g->ny * g->nx;
        sr = (size_t)a->n_shots * r->n;
        work[0] = a->work0;
        work[1] = a->work1;
        work[2] = a->work2;
        work[3] = a->work3;
        const ForwardPlan plan = adjoint_plan(g->ny, g->nx, a->work3, r);
        max_levels = plan.levels;
        wide = plan.wide;
        if (batched && a->hist_lap != 0)
            return fail(SWFD_EINVAL, "Laplacian history: not in batched runs");
        if (!wide) max_levels = std::min(max_levels, 2);
        recon = a->recon_next != nullptr || a->recon_cur != nullptr;
        if (max_levels >= 2)
            n_part = wide ? adjw_shape(g->ny, g->nx, a->n_shots, recon).n_groups
                          : adj2_shape(g->ny, g->nx, a->n_shots).n_groups;
        else
            n_part = image_groups(g->ny, g->nx, a->n_shots);
        lap = a->hist_lap != 0;
        if (lap && (!wide || max_levels < 2))
            return fail(SWFD_EINVAL, "Laplacian history needs full-width multi-step adjoint "
                                     "launches (see swfd_lap_history) and work3");
        if (recon) {
            if (a->recon_next == nullptr || a->recon_cur == nullptr || a->recon_work0 == nullptr ||
                a->recon_work1 == nullptr || a->recon_work2 == nullptr ||
                a->recon_work3 == nullptr || a->bnd == nullptr || a->bnd_slots < 1 ||
                a->src_rows == nullptr || a->src_cols == nullptr || a->src_values == nullptr)
                return fail(SWFD_EINVAL, "reconstruction: fields, work buffers, ring or sources "
                                         "missing");
            if (a->hist != nullptr || lap)
                return fail(SWFD_EINVAL, "reconstruction replaces the field history");
            if (a->grad == nullptr || a->workspace == nullptr)
                return fail(SWFD_EINVAL, "reconstruction: imaging needs grad and the workspace");
            if (!wide || max_levels < 2)
                return fail(SWFD_EINVAL, "reconstruction needs full-width adjoint launches (see "
                                         "swfd_adjoint_recon) and work3");
            rwork[0] = a->recon_work0;
            rwork[1] = a->recon_work1;
            rwork[2] = a->recon_work2;
            rwork[3] = a->recon_work3;
            r_next = a->recon_next;
            r_cur = a->recon_cur;
        }
        p_old = a->init_old;
        p_cur = a->init_u;
        n = a->step_hi - 1;
        return SWFD_OK;
    }
    bool done() const { return n < a->step_lo; }
    int next_levels() const
    {
        const int rem = n - a->step_lo + 1;
        int L = std::min(max_levels, rem);
        // reconstructing runs end with two or more steps where they can
        if (recon && L > 2 && rem - L == 1) --L;
        return L;
    }
    bool next_is_pair() const { return next_levels() >= 2; }
    bool next_is_wide() const { return wide && next_levels() >= 2; }
    long image_slot(int m) const
    {
        if (a->hist == nullptr || m < 1) return -1;
        const int j = m - 1;
        if (j < a->hist_first || (j - a->hist_first) % a->hist_every != 0) return -1;
        return (j - a->hist_first) / a->hist_every;
    }
    // Parameters of the next launch; `fill` is set when the imaging partials
    // must be cleared before it (first imaging launch of the run).
    int build(StepParams &p, bool &fill)
    {
        fill = false;
        base_params(p, grid, medium, a->n_shots);
        attach_receivers(p, rec);
        p.f_old = p_old;
        p.f_cur = p_cur;
        p.resid = a->residual + (size_t)n * sr;
        p.hist_lap = lap ? 1 : 0;
        auto use_partials = [&]() {
            if (!partials_used) fill = true;
            partials_used = true;
            p.g_part = a->workspace;
            p.img_scale = a->image_scale;
        };
        const int L = next_levels();
        p.levels = L;
        if (recon) {
            float *out_c = free_work(work, p_old, p_cur);
            float *out_d = free_work(work, p_old, p_cur, out_c);
            float *rout_c = free_work(rwork, r_next, r_cur);
            float *rout_d = free_work(rwork, r_next, r_cur, rout_c);
            if (out_c == nullptr || out_d == nullptr || rout_c == nullptr || rout_d == nullptr)
                return fail(SWFD_EINVAL, "adjoint: work buffers overlap state");
            for (int j = 0; j < L; ++j) p.lv_resid[j] = a->residual + (size_t)(n - j) * sr;
            const int lo = std::max(L - 2, 0);
            p.lv_out[lo] = out_c;
            p.lv_rout[lo] = rout_c;
            if (L >= 2) {
                p.lv_out[L - 1] = out_d;
                p.lv_rout[L - 1] = rout_d;
            }
            p.src_index = n;
            p.r_next = r_next;
            p.r_cur = r_cur;
            p.bnd = const_cast<float *>(a->bnd);
            p.bnd_first = a->bnd_first;
            p.bnd_slots = a->bnd_slots;
            p.src_rows = a->src_rows;
            p.src_cols = a->src_cols;
            p.src_values = a->src_values;
            p.src_w = a->src_weights;
            p.img_from = 1;  // step 0 images U(-1) = 0
            use_partials();
            return SWFD_OK;
        }
        if (L >= 2) {
            // level j is step n - j, imaged against the forward field before it
            float *out_c = free_work(work, p_old, p_cur);
            float *out_d = free_work(work, p_old, p_cur, out_c);
            if (out_c == nullptr || out_d == nullptr)
                return fail(SWFD_EINVAL, "adjoint: work buffers overlap state");
            bool any = false;
            for (int j = 0; j < L; ++j) {
                p.lv_resid[j] = a->residual + (size_t)(n - j) * sr;
                const long slot = image_slot(n - j);
                if (slot >= a->hist_slots)
                    return fail(SWFD_EINVAL, "adjoint: history slot out of range");
                if (slot >= 0) {
                    p.lv_in[j] = a->hist + (size_t)slot * field;
                    any = true;
                }
            }
            p.lv_out[L - 2] = out_c;
            p.lv_out[L - 1] = out_d;
            // two-step block launches
            p.f_out = out_c;
            p.f_out_b = out_d;
            p.resid_b = a->residual + (size_t)(n - 1) * sr;
            p.img = p.lv_in[0];
            p.img_b = p.lv_in[1];
            if (any) use_partials();
            return SWFD_OK;
        }
        float *out = free_work(work, p_old, p_cur);
        if (out == nullptr) return fail(SWFD_EINVAL, "adjoint: work buffers overlap state");
        p.f_out = out;
        if (max_levels < 2 && n < a->step_hi - 1) {
            const long slot = image_slot(n + 1);
            if (slot >= a->hist_slots)
                return fail(SWFD_EINVAL, "adjoint: history slot out of range");
            if (slot >= 0) {
                p.img = a->hist + (size_t)slot * field;
                use_partials();
            }
        }
        return SWFD_OK;
    }
    void advance(const StepParams &p)
    {
        const int L = p.levels;
        if (recon) {
            if (L >= 2) {
                p_old = p.lv_out[L - 2];
                p_cur = p.lv_out[L - 1];
                r_next = p.lv_rout[L - 2];
                r_cur = p.lv_rout[L - 1];
            } else {
                p_old = p_cur;
                p_cur = p.lv_out[0];
                r_next = r_cur;
                r_cur = p.lv_rout[0];
            }
            last_multi = true;  // every launch imaged its own steps
            n -= L;
            return;
        }
        if (L >= 2) {
            p_old = p.lv_out[L - 2];
            p_cur = p.lv_out[L - 1];
        } else {
            p_old = p_cur;
            p_cur = p.f_out;
        }
        last_multi = L >= 2;
        n -= L;
    }
    int launch_fill(hipStream_t st) const
    {
        const size_t nz = (size_t)n_part * plane;
        hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, st,
                           a->workspace, nz, 0.0f);
        SWFD_HIP_CHECK(hipGetLastError());
        return SWFD_OK;
    }
    int finish(hipStream_t st)
    {
        int rc;
        if (partials_used) {
            hipLaunchKernelGGL(reduce_groups_kernel, dim3((unsigned)((plane + 255) / 256)),
                               dim3(256), 0, st, plane, n_part, a->workspace, a->grad);
            SWFD_HIP_CHECK(hipGetLastError());
        }
        // the last step, unless a multi-step launch imaged it
        if (a->step_hi > a->step_lo && !last_multi) {
            const long slot = image_slot(a->step_lo);
            if (slot >= a->hist_slots)
                return fail(SWFD_EINVAL, "adjoint: history slot out of range");
            if (slot >= 0 &&
                (rc = launch_image(grid, a->n_shots, p_cur, a->hist + (size_t)slot * field,
                                   a->grad, a->grad, a->image_scale, st, lap)))
                return rc;
        }
        return settle_state(p_old, p_cur, a->final_old, a->final_u, work, field, st);
    }
};

// Batched tables: per tick, the StepParams of its entries and the first
// workgroup of each entry (n + 1 ints), each block aligned to 256 bytes.
constexpr size_t kTableAlign = 256;

size_t align_up(size_t v) { return (v + kTableAlign - 1) / kTableAlign * kTableAlign; }

// One launch of a tick: entries of one kernel variant (and workgroup size).
struct LaunchGroup {
    // 0: two-step blocks, 1: full width, 2: reconstructing adjoint (adjoint
    // batches) or receiver gathers (forward batches), 3: ring copies
    int kind = 0;
    unsigned flags = 0;
    int levels = 0, sm = 0, threads = 0;
    size_t lds = 0;
    bool image = false;  // two-step adjoint blocks: some entry images
    std::vector<StepParams> entries;
    std::vector<int> first{0};
    std::vector<int> runs;
    size_t bytes() const
    {
        return align_up(entries.size() * sizeof(StepParams)) +
               align_up((entries.size() + 1) * sizeof(int));
    }
};

// The ring copies of a run's next launch, appended to the tick's batched
// extraction group (entry: f_cur = field, bnd = ring slot, one thread per
// boundary cell of every shot).
int ForwardRun::boundary_entries(const StepParams &p, LaunchGroup &g) const
{
    const float *fields[kMaxLevels] = {nullptr, nullptr, nullptr, nullptr};
    float *slots[kMaxLevels] = {nullptr, nullptr, nullptr, nullptr};
    const int L = boundary_copies(p, fields, slots);
    if (L < 0) return -L;
    const size_t total = (size_t)a->n_shots * bnd_floats(grid->ny, grid->nx);
    for (int j = 0; j < L; ++j) {
        StepParams q;
        base_params(q, grid, nullptr, a->n_shots);
        q.f_cur = fields[j];
        q.bnd = slots[j];
        g.entries.push_back(q);
        g.first.push_back(g.first.back() + (int)((total + 255) / 256));
    }
    return SWFD_OK;
}

// The group of a tick for this variant (created in first-seen order, so the
// same in every mode).
LaunchGroup &group_for(std::vector<LaunchGroup> &groups, int kind, unsigned flags, int levels,
                       int sm, int threads)
{
    for (LaunchGroup &g : groups)
        if (g.kind == kind && g.flags == flags && g.levels == levels && g.sm == sm &&
            g.threads == threads)
            return g;
    groups.emplace_back();
    LaunchGroup &g = groups.back();
    g.kind = kind;
    g.flags = flags;
    g.levels = levels;
    g.sm = sm;
    g.threads = threads;
    return g;
}

// Mode 1: append the group's table to `host`; mode 2: launch it from the
// device copy at `table` + off (launch(group, table, first, n)).
template <typename Launch>
int emit_group(const LaunchGroup &g, int mode, std::vector<char> &host, size_t &off, void *table,
               size_t table_bytes, const char *what, Launch &&launch)
{
    const size_t nb = g.bytes();
    const size_t first_off = align_up(g.entries.size() * sizeof(StepParams));
    if (mode == 1) {
        host.resize(off + nb);
        std::memcpy(host.data() + off, g.entries.data(), g.entries.size() * sizeof(StepParams));
        std::memcpy(host.data() + off + first_off, g.first.data(), g.first.size() * sizeof(int));
    } else if (mode == 2) {
        if (off + nb > table_bytes)
            return fail(SWFD_EINVAL, std::string(what) + ": table buffer too small");
        const char *base = static_cast<const char *>(table) + off;
        int rc = launch(g, reinterpret_cast<const StepParams *>(base),
                        reinterpret_cast<const int *>(base + first_off), (int)g.entries.size());
        if (rc) return rc;
        SWFD_HIP_CHECK(hipGetLastError());
    }
    off += nb;
    return SWFD_OK;
}

int finish_tables(int mode, const std::vector<char> &host, size_t off, void *table,
                  size_t table_bytes, size_t *bytes_out, const char *what)
{
    if (bytes_out != nullptr) *bytes_out = off;
    if (mode == 1) {
        if (off > table_bytes) return fail(SWFD_EINVAL, std::string(what) + ": table buffer too small");
        if (off > 0) SWFD_HIP_CHECK(hipMemcpy(table, host.data(), off, hipMemcpyHostToDevice));
    }
    return SWFD_OK;
}

// mode 0: count table bytes; 1: write the tables to `table` (blocking copy);
// 2: launch (tables already in `table`). Modes 0 and 1 launch nothing. Each
// tick advances every unfinished run by one launch: the multi-step launches
// of the runs go out as one batched launch per kernel variant and row width,
// single steps one by one.
int forward_batch(int n_runs, const swfd_grid *grids, const swfd_medium *media,
                  const swfd_receivers *recs, const swfd_forward_args *args, int mode,
                  void *table, size_t table_bytes, size_t *bytes_out, void *stream)
{
    int rc;
    if (n_runs < 1 || grids == nullptr || media == nullptr || args == nullptr)
        return fail(SWFD_EINVAL, "forward batch: bad arguments");
    std::vector<ForwardRun> runs((size_t)n_runs);
    for (int i = 0; i < n_runs; ++i)
        if ((rc = runs[i].init(&grids[i], &media[i], recs != nullptr ? &recs[i] : nullptr,
                               &args[i], true)))
            return rc;
    hipStream_t st = as_stream(stream);
    std::vector<char> host;
    size_t off = 0;
    for (;;) {
        bool any = false;
        std::vector<LaunchGroup> groups;
        // boundary-ring copies of this tick's batched launches: one batched
        // extraction after them (the next tick's launches overwrite the fields)
        LaunchGroup ring;
        ring.kind = 3;
        ring.threads = 256;
        for (int i = 0; i < n_runs; ++i) {
            ForwardRun &r = runs[i];
            if (r.done()) continue;
            any = true;
            StepParams p;
            const int gathers = r.build(p);
            if (r.next_is_wide()) {
                int threads = 0, sm = 0;
                unsigned f = 0;
                const int blocks = multi_shape(p, gathers, threads, f, sm);
                if (blocks < 0) return -blocks;
                LaunchGroup &g = group_for(groups, 1, f, p.levels, sm, threads);
                g.entries.push_back(p);
                g.first.push_back(g.first.back() + blocks);
                g.runs.push_back(i);
                if ((rc = r.boundary_entries(p, ring))) return rc;
            } else if (r.next_is_pair()) {
                const int blocks = pair_geometry(p, gathers);
                LaunchGroup &g = group_for(groups, 0, pair_flags(p, gathers > 0), 2, 0,
                                           kLanes * kWavesPerBlock);
                g.entries.push_back(p);
                g.first.push_back(g.first.back() + blocks);
                g.runs.push_back(i);
                if ((rc = r.boundary_entries(p, ring))) return rc;
            } else {
                if (mode == 2 && ((rc = launch_step(p, false, st, gathers)) ||
                                  (rc = r.launch_gather(p, st)) ||
                                  (rc = r.launch_boundary(p, st))))
                    return rc;
                r.advance(p);
            }
        }
        if (!any) break;
        // receiver samples of every level of this tick's full-width launches, for
        // the runs that gather after their launches: one batched launch after them
        LaunchGroup gather;
        gather.kind = 2;
        gather.threads = kGatherThreads;
        for (const LaunchGroup &g : groups)
            for (size_t j = 0; j < g.runs.size(); ++j) {
                const ForwardRun &r = runs[g.runs[j]];
                if (!r.sep_gather || g.kind != 1) continue;
                const StepParams &p = g.entries[j];
                const int pairs = p.levels * p.n_shots * p.n_rec;
                gather.entries.push_back(p);
                gather.first.push_back(gather.first.back() +
                                       (pairs + kGatherThreads - 1) / kGatherThreads);
            }
        for (const LaunchGroup &g : groups) {
            rc = emit_group(g, mode, host, off, table, table_bytes, "forward batch",
                            [&](const LaunchGroup &gg, const StepParams *tab, const int *first,
                                int ne) {
                                const dim3 grid((unsigned)gg.first.back()), block(gg.threads);
                                if (gg.kind == 0) {
                                    launch_fwd2_variant(gg.flags, grid, st, [&](auto, auto batched) {
                                        hipLaunchKernelGGL(batched, grid, block, 0, st, tab, first, ne);
                                    });
                                    return SWFD_OK;
                                }
                                return fwdm_levels_variant(gg.levels, gg.sm, gg.flags,
                                                           [&](auto, auto batched) {
                                    hipLaunchKernelGGL(batched, grid, block, 0, st, tab, first, ne);
                                });
                            });
            if (rc) return rc;
            for (size_t j = 0; j < g.runs.size(); ++j) runs[g.runs[j]].advance(g.entries[j]);
        }
        if (!gather.entries.empty()) {
            rc = emit_group(gather, mode, host, off, table, table_bytes, "forward batch",
                            [&](const LaunchGroup &gg, const StepParams *tab, const int *first,
                                int ne) {
                                hipLaunchKernelGGL(gather_levels_batch_kernel,
                                                   dim3((unsigned)gg.first.back()),
                                                   dim3(kGatherThreads), 0, st, tab, first, ne);
                                return SWFD_OK;
                            });
            if (rc) return rc;
        }
        if (!ring.entries.empty()) {
            rc = emit_group(ring, mode, host, off, table, table_bytes, "forward batch",
                            [&](const LaunchGroup &gg, const StepParams *tab, const int *first,
                                int ne) {
                                hipLaunchKernelGGL(bnd_extract_batch_kernel,
                                                   dim3((unsigned)gg.first.back()), dim3(256), 0,
                                                   st, tab, first, ne);
                                return SWFD_OK;
                            });
            if (rc) return rc;
        }
    }
    if ((rc = finish_tables(mode, host, off, table, table_bytes, bytes_out, "forward batch")))
        return rc;
    if (mode == 2)
        for (ForwardRun &r : runs)
            if ((rc = r.finish(stream))) return rc;
    return SWFD_OK;
}

int adjoint_batch(int n_runs, const swfd_grid *grids, const swfd_medium *media,
                  const swfd_receivers *recs, const swfd_adjoint_args *args, int mode,
                  void *table, size_t table_bytes, size_t *bytes_out, void *stream)
{
    int rc;
    if (n_runs < 1 || grids == nullptr || media == nullptr || recs == nullptr || args == nullptr)
        return fail(SWFD_EINVAL, "adjoint batch: bad arguments");
    std::vector<AdjointRun> runs((size_t)n_runs);
    for (int i = 0; i < n_runs; ++i)
        if ((rc = runs[i].init(&grids[i], &media[i], &recs[i], &args[i], true))) return rc;
    hipStream_t st = as_stream(stream);
    std::vector<char> host;
    size_t off = 0;
    for (;;) {
        bool any = false;
        std::vector<LaunchGroup> groups;
        for (int i = 0; i < n_runs; ++i) {
            AdjointRun &r = runs[i];
            if (r.done()) continue;
            any = true;
            StepParams p;
            bool fill = false;
            if ((rc = r.build(p, fill))) return rc;
            if (mode == 2 && fill && (rc = r.launch_fill(st))) return rc;
            if (r.recon) {
                int threads = 0;
                size_t lds = 0;
                const int blocks = adjr_shape(p, threads, lds);
                if (blocks < 0) return -blocks;
                adjr_set_lean(p);
                LaunchGroup &g = group_for(groups, 2, adjr_every(p) ? 1u : 0u, p.levels, 0, threads);
                g.lds = std::max(g.lds, lds);
                g.entries.push_back(p);
                g.first.push_back(g.first.back() + blocks);
                g.runs.push_back(i);
            } else if (r.next_is_wide()) {
                int threads = 0;
                size_t lds = 0;
                const int blocks = adjm_shape(p, threads, lds);
                if (blocks < 0) return -blocks;
                LaunchGroup &g = group_for(groups, 1, (unsigned)adjm_image(p), p.levels, 0, threads);
                g.lds = std::max(g.lds, lds);
                g.entries.push_back(p);
                g.first.push_back(g.first.back() + blocks);
                g.runs.push_back(i);
            } else if (r.next_is_pair()) {
                size_t lds = 0;
                const int blocks = adj_pair_geometry(p, lds);
                // the batched launch runs every entry with the imaging variant
                // when any entry images, so size the LDS for every entry
                const size_t acc = (size_t)kWavesPerBlock * p.acc_rows * kLanes * sizeof(float);
                if (acc > 64 * 1024)
                    return fail(SWFD_EINVAL, "two-step adjoint: row segments too long");
                LaunchGroup &g = group_for(groups, 0, 0, 2, 0, kLanes * kWavesPerBlock);
                if (lds > 0) g.image = true;
                g.lds = std::max(g.lds, acc);
                g.entries.push_back(p);
                g.first.push_back(g.first.back() + blocks);
                g.runs.push_back(i);
            } else {
                if (mode == 2 && (rc = launch_step(p, true, st, 0))) return rc;
                r.advance(p);
            }
        }
        if (!any) break;
        for (const LaunchGroup &g : groups) {
            rc = emit_group(g, mode, host, off, table, table_bytes, "adjoint batch",
                            [&](const LaunchGroup &gg, const StepParams *tab, const int *first,
                                int ne) {
                                const dim3 grid((unsigned)gg.first.back()), block(gg.threads);
                                if (gg.kind == 2)
                                    return adjr_variant(gg.levels, gg.flags != 0,
                                                        [&](auto, auto batched) {
                                        hipLaunchKernelGGL(batched, grid, block, gg.lds, st, tab,
                                                           first, ne);
                                    });
                                if (gg.kind == 0) {
                                    if (gg.image)
                                        hipLaunchKernelGGL((adj2_batch_kernel<kResid | kImage>), grid,
                                                           block, gg.lds, st, tab, first, ne);
                                    else
                                        hipLaunchKernelGGL((adj2_batch_kernel<kResid>), grid, block,
                                                           0, st, tab, first, ne);
                                    return SWFD_OK;
                                }
                                return adjm_levels_variant(gg.levels, (int)gg.flags,
                                                           [&](auto, auto batched) {
                                    hipLaunchKernelGGL(batched, grid, block, gg.lds, st, tab, first,
                                                       ne);
                                });
                            });
            if (rc) return rc;
            for (size_t j = 0; j < g.runs.size(); ++j) runs[g.runs[j]].advance(g.entries[j]);
        }
    }
    if ((rc = finish_tables(mode, host, off, table, table_bytes, bytes_out, "adjoint batch")))
        return rc;
    if (mode == 2)
        for (AdjointRun &r : runs)
            if ((rc = r.finish(st))) return rc;
    return SWFD_OK;
}

}  // namespace

// Partial misfit sums of swfd_residual_noalloc: one float per residual tile,
// SWFD_RESIDUAL_SCRATCH_FLOATS of them (4 MiB). Every GPU the library runs on
// holds its own copy until the process ends. Host code never uses this name
// as an address: it takes each device's address from hipGetSymbolAddress.
static __device__ __align__(256) float scratch[SWFD_RESIDUAL_SCRATCH_FLOATS];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
float *g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
int scratch_on_current_device(float **out)
{
    int dev = -1;
    SWFD_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices)
        return fail(SWFD_EINVAL, "residual (no allocation): device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        SWFD_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) return fail(SWFD_EHIP, "residual (no allocation): no scratch address");
        g_scratch_ptr[dev] = static_cast<float *>(p);
    }
    *out = g_scratch_ptr[dev];
    return SWFD_OK;
}

int check_residual_args(int32_t n_steps, int32_t n_shots, int32_t n_rec, const float *samples,
                        const float *observed, const int32_t *obs_index, const int32_t *mute,
                        const float *residual, const float *misfit)
{
    if (n_steps < 1 || n_shots < 1 || n_rec < 1)
        return fail(SWFD_EINVAL, "residual: sizes must be positive");
    if (samples == nullptr || observed == nullptr || obs_index == nullptr ||
        mute == nullptr || residual == nullptr || misfit == nullptr)
        return fail(SWFD_EINVAL, "residual: NULL pointer");
    return SWFD_OK;
}

// residual_kernel over every tile (one partial per tile into `partials`), then
// misfit_kernel over those partials; both write before they read, so the
// partials need no clearing.
int launch_residual(int32_t n_steps, int32_t n_shots, int32_t n_rec, const float *samples,
                    const float *observed, int64_t obs_stride_e, int64_t obs_stride_r,
                    const int32_t *obs_index, const int32_t *mute, float *residual,
                    float *partials, size_t tiles, float *misfit, hipStream_t st)
{
    hipLaunchKernelGGL(residual_kernel, dim3((unsigned)tiles), dim3(256), 0, st, n_steps,
                       n_shots, n_rec, samples, observed, (long long)obs_stride_e,
                       (long long)obs_stride_r, obs_index, mute, residual, partials);
    SWFD_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(misfit_kernel, dim3(1), dim3(256), 0, st, (int)tiles, partials, misfit);
    SWFD_HIP_CHECK(hipGetLastError());
    return SWFD_OK;
}

}  // namespace

extern "C" {

const char *swfd_last_error(void) { return g_last_error.c_str(); }

const char *swfd_version(void) { return "swfd 0.1 gfx950"; }

size_t swfd_medium_floats(int32_t ny, int32_t nx)
{
    return 2 * (size_t)ny * nx + 2 * (size_t)ny + 2 * (size_t)nx + 8;
}

int swfd_footprint_rows(int32_t ny, int32_t nx, int32_t n, const int32_t *row0,
                        const int32_t *col0, uint32_t *out)
{
    if (ny < 3 || nx < 3 || n < 0 || (n > 0 && (row0 == nullptr || col0 == nullptr)) ||
        out == nullptr)
        return fail(SWFD_EINVAL, "footprint rows: bad arguments");
    for (int y = 0; y < ny; ++y) out[y] = 0u;
    if (!wide_fits(nx)) return SWFD_OK;
    // the column layout of full-width launches (wide_geometry / adjoint shape)
    const int off = (nx % kLanes == 1) ? 1 : 0;
    for (int i = 0; i < n; ++i) {
        if (row0[i] < 0 || row0[i] + 1 >= ny || col0[i] < 0 || col0[i] + 1 >= nx)
            return fail(SWFD_EINVAL, "footprint rows: receiver footprint outside the grid");
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx)
                out[row0[i] + dy] |= 1u << ((col0[i] + dx + off) / kLanes);
    }
    return SWFD_OK;
}

void swfd_medium_view(const swfd_grid *grid, const float *coef, swfd_medium *out)
{
    const size_t n = (size_t)grid->ny * grid->nx;
    out->c2 = coef;
    out->c2dt2 = coef + n;
    out->mur = coef + 2 * n;
}

int swfd_prepare_medium(const swfd_grid *grid, const float *c, double dt, double dx,
                        double dy, float *coef, void *stream)
{
    int rc;
    if ((rc = check_grid(grid))) return rc;
    if (c == nullptr || coef == nullptr) return fail(SWFD_EINVAL, "c / coef is NULL");
    const size_t n = (size_t)grid->ny * grid->nx;
    const size_t work = std::max<size_t>(n, (size_t)std::max(grid->ny, grid->nx));
    const unsigned blocks = (unsigned)((work + 255) / 256);
    hipLaunchKernelGGL(prepare_medium_kernel, dim3(blocks), dim3(256), 0, as_stream(stream),
                       grid->ny, grid->nx, c, (float)dt, (float)dx, (float)dy, grid->dt2,
                       coef, coef + n, coef + 2 * n);
    SWFD_HIP_CHECK(hipGetLastError());
    return SWFD_OK;
}

int swfd_step(const swfd_grid *grid, const swfd_medium *medium, const swfd_receivers *rec,
              const swfd_step_args *a, void *stream)
{
    int rc;
    if ((rc = check_grid(grid)) || (rc = check_medium(medium))) return rc;
    if (a == nullptr) return fail(SWFD_EINVAL, "step args are NULL");
    if (a->n_shots < 1) return fail(SWFD_EINVAL, "n_shots must be positive");
    if (a->u_old == nullptr || a->u == nullptr || a->u_next == nullptr)
        return fail(SWFD_EINVAL, "field pointers are NULL");
    if (a->u_next == a->u_old || a->u_next == a->u)
        return fail(SWFD_EINVAL, "u_next must not alias u_old or u");
    if (a->src_rows != nullptr && (a->src_cols == nullptr || a->src_values == nullptr))
        return fail(SWFD_EINVAL, "point source needs rows, cols and values");
    if (a->residual != nullptr && (rc = check_receivers(rec, true))) return rc;
    if (a->samples != nullptr && (rc = check_receivers(rec, false))) return rc;
    if (a->image_field != nullptr && (a->grad_in == nullptr || a->grad_out == nullptr))
        return fail(SWFD_EINVAL, "imaging needs grad_in and grad_out");
    if (a->image_field != nullptr && !a->adjoint)
        return fail(SWFD_EINVAL, "imaging is an adjoint-step option");
    StepParams p;
    base_params(p, grid, medium, a->n_shots);
    attach_receivers(p, a->residual != nullptr ? rec : nullptr);
    p.f_old = a->u_old;
    p.f_cur = a->u;
    p.f_out = a->u_next;
    p.f_out2 = a->u_next_copy;
    p.src_rows = a->src_rows;
    p.src_cols = a->src_cols;
    p.src_values = a->src_values;
    p.src_index = a->src_index;
    p.src_w = a->src_rows != nullptr ? a->src_weights : nullptr;
    p.dense = a->dense;
    p.resid = a->residual;
    if ((rc = launch_step(p, a->adjoint != 0, as_stream(stream), 0))) return rc;
    if (a->image_field != nullptr &&
        (rc = launch_image(grid, a->n_shots, a->u_next, a->image_field, a->grad_in,
                           a->grad_out, a->image_scale, as_stream(stream))))
        return rc;
    if (a->samples != nullptr)
        return swfd_gather(grid, rec, a->n_shots, a->u_next, a->samples, stream);
    return SWFD_OK;
}

int swfd_gather(const swfd_grid *grid, const swfd_receivers *rec, int32_t n_shots,
                const float *fields, float *samples, void *stream)
{
    int rc;
    if ((rc = check_grid(grid)) || (rc = check_receivers(rec, false))) return rc;
    if (fields == nullptr || samples == nullptr || n_shots < 1)
        return fail(SWFD_EINVAL, "gather: bad fields / samples / n_shots");
    StepParams p;
    base_params(p, grid, nullptr, n_shots);
    attach_receivers(p, rec);
    p.gat_field = fields;
    p.gat_out = samples;
    const int pairs = n_shots * rec->n;
    hipLaunchKernelGGL(gather_kernel, dim3((pairs + 255) / 256), dim3(256), 0,
                       as_stream(stream), p);
    SWFD_HIP_CHECK(hipGetLastError());
    return SWFD_OK;
}

int swfd_scatter(const swfd_grid *grid, const swfd_receivers *rec, int32_t n_shots,
                 const float *residual, float *fields, void *stream)
{
    int rc;
    if ((rc = check_grid(grid)) || (rc = check_receivers(rec, true))) return rc;
    if (residual == nullptr || fields == nullptr || n_shots < 1)
        return fail(SWFD_EINVAL, "scatter: bad residual / fields / n_shots");
    const size_t plane = (size_t)grid->ny * grid->nx;
    const size_t total = plane * n_shots;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       as_stream(stream), fields, total, 0.0f);
    SWFD_HIP_CHECK(hipGetLastError());
    const int items = n_shots * rec->n_entries;
    if (items > 0) {
        hipLaunchKernelGGL(scatter_kernel, dim3((items + 255) / 256), dim3(256), 0,
                           as_stream(stream), n_shots, plane, rec->n_entries, rec->n,
                           residual, rec->entry_cell, rec->entry_start, rec->pair_rec,
                           rec->pair_w, fields);
        SWFD_HIP_CHECK(hipGetLastError());
    }
    return SWFD_OK;
}

int swfd_laplacian(const swfd_grid *grid, int32_t n_fields, const float *fields, float *out,
                   void *stream)
{
    int rc;
    if ((rc = check_grid(grid))) return rc;
    if (fields == nullptr || out == nullptr || n_fields < 1 || fields == out)
        return fail(SWFD_EINVAL, "laplacian: bad fields / out / count");
    const size_t total = (size_t)grid->ny * grid->nx * n_fields;
    hipLaunchKernelGGL(laplacian_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       as_stream(stream), grid->ny, grid->nx, total, grid->inv_dx2,
                       grid->inv_dy2, grid->center, fields, out);
    SWFD_HIP_CHECK(hipGetLastError());
    return SWFD_OK;
}

int swfd_history_convert(size_t n, int32_t direction, int32_t type, const void *src, void *dst,
                         void *stream)
{
    if (type != SWFD_HISTORY_F16 && type != SWFD_HISTORY_BF16)
        return fail(SWFD_EINVAL, "history convert: type must be float16 or bfloat16");
    if (direction != 0 && direction != 1)
        return fail(SWFD_EINVAL, "history convert: direction must be 0 (pack) or 1 (unpack)");
    if (n == 0) return SWFD_OK;
    if (src == nullptr || dst == nullptr || src == dst)
        return fail(SWFD_EINVAL, "history convert: bad src / dst");
    const bool vec = n % 8 == 0 && ((uintptr_t)src % 16) == 0 && ((uintptr_t)dst % 16) == 0;
    const size_t lanes = vec ? n / 8 : n;
    const size_t blocks = (lanes + 255) / 256;
    if (blocks > 0x7fffffffu) return fail(SWFD_EINVAL, "history convert: too many values");
    hipStream_t st = as_stream(stream);
    const dim3 g((unsigned)blocks), b(256);
    if (direction == 0) {
        if (vec) hipLaunchKernelGGL((history_convert_kernel<true, 8>), g, b, 0, st, n, src, dst, type);
        else hipLaunchKernelGGL((history_convert_kernel<true, 1>), g, b, 0, st, n, src, dst, type);
    } else {
        if (vec) hipLaunchKernelGGL((history_convert_kernel<false, 8>), g, b, 0, st, n, src, dst, type);
        else hipLaunchKernelGGL((history_convert_kernel<false, 1>), g, b, 0, st, n, src, dst, type);
    }
    SWFD_HIP_CHECK(hipGetLastError());
    return SWFD_OK;
}

int swfd_forward(const swfd_grid *grid, const swfd_medium *medium, const swfd_receivers *rec,
                 const swfd_forward_args *a, void *stream)
{
    int rc;
    ForwardRun run;
    if ((rc = run.init(grid, medium, rec, a))) return rc;
    hipStream_t st = as_stream(stream);
    while (!run.done()) {
        StepParams p;
        const int gathers = run.build(p);
        if (run.next_is_wide()) rc = launch_multi(p, st, gathers);
        else if (run.next_is_pair()) rc = launch_pair(p, st, gathers);
        else rc = launch_step(p, false, st, gathers);
        if (rc || (rc = run.launch_gather(p, st)) || (rc = run.launch_boundary(p, st))) return rc;
        run.advance(p);
    }
    return run.finish(stream);
}

int swfd_adjoint(const swfd_grid *grid, const swfd_medium *medium, const swfd_receivers *rec,
                 const swfd_adjoint_args *a, void *stream)
{
    int rc;
    AdjointRun run;
    if ((rc = run.init(grid, medium, rec, a))) return rc;
    hipStream_t st = as_stream(stream);
    while (!run.done()) {
        StepParams p;
        bool fill = false;
        if ((rc = run.build(p, fill))) return rc;
        if (fill && (rc = run.launch_fill(st))) return rc;
        if (run.recon) rc = launch_adj_recon(p, st);
        else if (run.next_is_wide()) rc = launch_adj_multi(p, st);
        else if (run.next_is_pair()) rc = launch_adj_pair(p, st);
        else rc = launch_step(p, true, st, 0);
        if (rc) return rc;
        run.advance(p);
    }
    return run.finish(st);
}

int swfd_forward_batch_table_bytes(int32_t n, const swfd_grid *grids, const swfd_medium *media,
                                   const swfd_receivers *recs, const swfd_forward_args *args,
                                   size_t *bytes)
{
    return forward_batch(n, grids, media, recs, args, 0, nullptr, 0, bytes, nullptr);
}

int swfd_forward_batch_prepare(int32_t n, const swfd_grid *grids, const swfd_medium *media,
                               const swfd_receivers *recs, const swfd_forward_args *args,
                               void *table, size_t table_bytes)
{
    return forward_batch(n, grids, media, recs, args, 1, table, table_bytes, nullptr, nullptr);
}

int swfd_forward_batch(int32_t n, const swfd_grid *grids, const swfd_medium *media,
                       const swfd_receivers *recs, const swfd_forward_args *args,
                       const void *table, size_t table_bytes, void *stream)
{
    return forward_batch(n, grids, media, recs, args, 2, const_cast<void *>(table), table_bytes,
                         nullptr, stream);
}

int swfd_adjoint_batch_table_bytes(int32_t n, const swfd_grid *grids, const swfd_medium *media,
                                   const swfd_receivers *recs, const swfd_adjoint_args *args,
                                   size_t *bytes)
{
    return adjoint_batch(n, grids, media, recs, args, 0, nullptr, 0, bytes, nullptr);
}

int swfd_adjoint_batch_prepare(int32_t n, const swfd_grid *grids, const swfd_medium *media,
                               const swfd_receivers *recs, const swfd_adjoint_args *args,
                               void *table, size_t table_bytes)
{
    return adjoint_batch(n, grids, media, recs, args, 1, table, table_bytes, nullptr, nullptr);
}

int swfd_adjoint_batch(int32_t n, const swfd_grid *grids, const swfd_medium *media,
                       const swfd_receivers *recs, const swfd_adjoint_args *args,
                       const void *table, size_t table_bytes, void *stream)
{
    return adjoint_batch(n, grids, media, recs, args, 2, const_cast<void *>(table), table_bytes,
                         nullptr, stream);
}

int swfd_tune(int32_t kind, int32_t ny, int32_t nx, int32_t seg_rows, int32_t param,
              int32_t depth)
{
    if (kind < 0 || kind > 3 || ny < 3 || nx < 3)
        return fail(SWFD_EINVAL, "tune: kind must be 0..3 and the grid at least 3 x 3");
    for (TuneEntry &t : g_tune)
        if (t.kind == kind && t.ny == ny && t.nx == nx) {
            t.seg_rows = seg_rows;
            t.param = param;
            t.depth = depth;
            return SWFD_OK;
        }
    g_tune.push_back(TuneEntry{kind, ny, nx, seg_rows, param, depth});
    return SWFD_OK;
}

void swfd_tune_clear(void) { g_tune.clear(); }

int32_t swfd_steps_per_launch(int32_t kind, int32_t ny, int32_t nx, int32_t samples)
{
    if ((kind != 0 && kind != 1) || ny < 3 || nx < 3) return 0;
    static const int32_t any_map = 0;
    swfd_receivers rec{};
    rec.cell_map = &any_map;
    static const float any_buffer = 0.0f;
    float *w3 = const_cast<float *>(&any_buffer);
    if (kind == 1) return adjoint_plan(ny, nx, w3, &rec).levels;
    const ForwardPlan f = forward_plan(ny, nx, w3);
    int levels = f.levels;
    if (!f.wide || samples) levels = std::min(levels, 2);
    return levels;
}

int swfd_adjoint_launch_shape(int32_t ny, int32_t nx, int32_t n_shots, int32_t recon,
                              int32_t *out)
{
    if (ny < 3 || nx < 3 || n_shots < 1 || out == nullptr)
        return fail(SWFD_EINVAL, "adjoint launch shape: bad arguments");
    static const int32_t any_map = 0;
    swfd_receivers rec{};
    rec.cell_map = &any_map;
    static const float any_buffer = 0.0f;
    float *w3 = const_cast<float *>(&any_buffer);
    const ForwardPlan plan = adjoint_plan(ny, nx, w3, &rec);
    for (int i = 0; i < 6; ++i) out[i] = 0;
    out[0] = plan.levels;
    out[1] = plan.wide ? 1 : 0;
    if (plan.wide && plan.levels >= 2) {
        const AdjwShape a = adjw_shape(ny, nx, n_shots, recon != 0);
        out[2] = a.acc_rows;
        out[3] = a.n_seg;
        out[4] = a.n_groups;
        out[5] = a.shots_per_group;
    }
    return SWFD_OK;
}

size_t swfd_boundary_floats(int32_t ny, int32_t nx)
{
    if (ny < 3 || nx < 3) return 0;
    return bnd_floats(ny, nx);
}

int32_t swfd_adjoint_recon(int32_t ny, int32_t nx)
{
    if (ny < 3 || nx < 3) return 0;
    static const int32_t any_map = 0;
    swfd_receivers rec{};
    rec.cell_map = &any_map;
    static const float any_buffer = 0.0f;
    float *w3 = const_cast<float *>(&any_buffer);
    const ForwardPlan a = adjoint_plan(ny, nx, w3, &rec);
    return (a.wide && a.levels >= 2) ? 1 : 0;
}

int32_t swfd_lap_history(int32_t ny, int32_t nx)
{
    if (ny < 3 || nx < 3) return 0;
    static const int32_t any_map = 0;
    swfd_receivers rec{};
    rec.cell_map = &any_map;  // the adjoint's full-width launches need a scatter table
    static const float any_buffer = 0.0f;
    float *w3 = const_cast<float *>(&any_buffer);
    const ForwardPlan f = forward_plan(ny, nx, w3), a = adjoint_plan(ny, nx, w3, &rec);
    return (f.wide && f.levels >= 2 && a.wide && a.levels >= 2) ? 1 : 0;
}

size_t swfd_adjoint_workspace_floats(int32_t ny, int32_t nx, int32_t n_shots)
{
    const int groups = std::max(kMaxImageGroups, std::min(max_wide_groups(ny, nx), (int)n_shots));
    return (size_t)groups * (size_t)ny * (size_t)nx;
}

size_t swfd_residual_workspace_floats(int32_t n_steps, int32_t n_shots, int32_t n_rec)
{
    const size_t tiles_r = (n_rec + kResTile - 1) / kResTile;
    const size_t tiles_n = (n_steps + kResTile - 1) / kResTile;
    return tiles_r * tiles_n * (size_t)n_shots;
}

int swfd_residual(int32_t n_steps, int32_t n_shots, int32_t n_rec, const float *samples,
                  const float *observed, int64_t obs_stride_e, int64_t obs_stride_r,
                  const int32_t *obs_index, const int32_t *mute, float *residual,
                  float *workspace, float *misfit, void *stream)
{
    int rc = check_residual_args(n_steps, n_shots, n_rec, samples, observed, obs_index, mute,
                                 residual, misfit);
    if (rc) return rc;
    if (workspace == nullptr) return fail(SWFD_EINVAL, "residual: NULL pointer");
    const size_t blocks = swfd_residual_workspace_floats(n_steps, n_shots, n_rec);
    if (blocks > 0x7fffffff) return fail(SWFD_EINVAL, "residual: too many tiles");
    return launch_residual(n_steps, n_shots, n_rec, samples, observed, obs_stride_e,
                           obs_stride_r, obs_index, mute, residual, workspace, blocks, misfit,
                           as_stream(stream));
}

int swfd_residual_noalloc(int32_t n_steps, int32_t n_shots, int32_t n_rec, const float *samples,
                          const float *observed, int64_t obs_stride_e, int64_t obs_stride_r,
                          const int32_t *obs_index, const int32_t *mute, float *residual,
                          float *misfit, void *stream)
{
    int rc = check_residual_args(n_steps, n_shots, n_rec, samples, observed, obs_index, mute,
                                 residual, misfit);
    if (rc) return rc;
    const size_t tiles = swfd_residual_workspace_floats(n_steps, n_shots, n_rec);
    if (tiles > (size_t)SWFD_RESIDUAL_SCRATCH_FLOATS)
        return fail(SWFD_EINVAL,
                    "residual (no allocation): " + std::to_string(tiles) +
                        " partial sums needed, the built-in scratch holds " +
                        std::to_string((size_t)SWFD_RESIDUAL_SCRATCH_FLOATS) +
                        " (64 x 64 tiles of steps x receivers per shot); use swfd_residual "
                        "with a workspace");
    float *partials = nullptr;
    if ((rc = scratch_on_current_device(&partials))) return rc;
    return launch_residual(n_steps, n_shots, n_rec, samples, observed, obs_stride_e,
                           obs_stride_r, obs_index, mute, residual, partials, tiles, misfit,
                           as_stream(stream));
}

}  // extern "C"
