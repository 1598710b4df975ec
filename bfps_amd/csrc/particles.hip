// Tracer interpolation and Adams-Bashforth update (gfx950).
//
// Reference: particles_field_computer::apply_computation
// (bfps/cpp/particles/particles_field_computer.hpp:55-183) on one rank,
// particles_adams_bashforth::move_particles (particles_adams_bashforth.hpp:18-133),
// particles_system::completeLoop (particles_system.hpp:173-217).
//
// Tracer data stay in index order in device memory.  Each loop first sorts a
// permutation of the tracers by the Morton code of their 4^3-cell
// micro-brick (the reference buckets by z-slab for the same reason,
// particles_utils.hpp:106-165), so consecutive threads interpolate nearby
// tracers and share stencil rows in L1/L2; blocks are mapped so that each
// XCD receives a contiguous run of the sorted order.
//
// Per tracer and axis: cell = floor(pos / h) mod N, x = position in the cell
// in [0, 1); weights beta_j(x), j = 0..2n+1, for nodes cell-n .. cell+n+1
// (wrapped).  The field sum is taken row by row along x (contiguous in the
// [y][z][x] real layout): acc += (bz * by) * sum_x f * bx, in double; the
// reference sums f * ((bz * by) * bx) point by point, which differs only by
// rounding.
#include <hipcub/hipcub.hpp>

#include <cstdio>
#include <cstdlib>

#include "ctx.hpp"

namespace {

struct PartGeomD {
    int n;
    double width[3];  // box widths 2 pi / dk (particles_system_builder.hpp:175-177)
    double h[3];      // grid spacings width / N (:187-189)
};

PartGeomD pgeom_d(const bfpsh_ctx* ctx) {
    PartGeomD p;
    p.n = ctx->n;
    const double dkv[3] = {ctx->dkx, ctx->dky, ctx->dkz};
    for (int d = 0; d < 3; d++) {
        p.width[d] = 4 * std::acos(0.0) / dkv[d];
        p.h[d] = p.width[d] / double(ctx->n);
    }
    return p;
}

// pbc_field_layer (particles_field_computer.hpp:176-183)
__device__ __forceinline__ int cell_of(double s, double h, int n) {
    const int lvl = int(floor(s / h));
    return ((lvl % n) + n) % n;
}

__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_part_keys(const double* __restrict__ pos, int64_t np, PartGeomD pg, int shift,
                            uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int cx = cell_of(pos[p * 3 + 0], pg.h[0], pg.n) >> shift;
    const int cy = cell_of(pos[p * 3 + 1], pg.h[1], pg.n) >> shift;
    const int cz = cell_of(pos[p * 3 + 2], pg.h[2], pg.n) >> shift;
    keys[p] = spread3(cx) | (spread3(cy) << 1) | (spread3(cz) << 2);
    vals[p] = uint32_t(p);
}

// logical block index so that XCD k (blocks dispatched round-robin over 8
// XCDs) processes the k-th contiguous eighth of the sorted order
__device__ __forceinline__ int64_t xcd_block() {
    const int64_t nb8 = gridDim.x / 8;
    const int64_t b = blockIdx.x;
    return (b % 8) * nb8 + b / 8;
}

template <int NB>
__device__ __forceinline__ void setup_axis(double s, const double* __restrict__ tab, int deg, double width, double h,
                                           int n, double (&bw)[2 * NB + 2], int (&idx)[2 * NB + 2]) {
    constexpr int W = 2 * NB + 2;
    // get_norm_pos_in_cell (particles_field_computer.hpp:55-63)
    const double rep = floor(s / width);
    const double inbox = s - rep * width;
    const double ci = floor(inbox / h);
    const double xr = (inbox - ci * h) / h;
#pragma unroll
    for (int j = 0; j < W; j++) {
        double acc = 0.0;
        for (int q = deg; q >= 0; q--) acc = acc * xr + tab[j * (deg + 1) + q];
        bw[j] = acc;
    }
    const int c = cell_of(s, h, n);
#pragma unroll
    for (int j = 0; j < W; j++) {
        int v = c - NB + j;
        v = v < 0 ? v + n : (v >= n ? v - n : v);
        idx[j] = v;
    }
}

// Lanes per tracer: the W = 2n+2 stencil columns along x are spread over LP
// lanes (the power of two from 4 to 32 that holds W), so one wavefront load
// instruction reads 64/LP contiguous row segments instead of 64 scattered
// words, with few idle lanes for small stencils.
template <int NB> struct Lanes {
    static constexpr int W = 2 * NB + 2;
    static constexpr int LP = W <= 4 ? 4 : (W <= 8 ? 8 : (W <= 16 ? 16 : 32));
};

// sum over the stencil of one tracer, computed by the LP lanes of its group;
// lane jx < W holds column jx.  Result valid in every lane of the group.
template <typename T, int NB, int NC>
__device__ __forceinline__ void gather_coop(const T* __restrict__ f, int n, const double* x,
                                            const double* __restrict__ tab, int deg, const PartGeomD& pg,
                                            int jx, double (&acc)[NC]) {
    constexpr int W = 2 * NB + 2;
    constexpr int LP = Lanes<NB>::LP;
    double by[W], bz[W];
    int iy[W], iz[W];
    setup_axis<NB>(x[1], tab, deg, pg.width[1], pg.h[1], n, by, iy);
    setup_axis<NB>(x[2], tab, deg, pg.width[2], pg.h[2], n, bz, iz);
    // this lane's x column
    double bxl = 0.0;
    int ixl = 0;
    {
        const double s = x[0];
        const double rep = floor(s / pg.width[0]);
        const double inbox = s - rep * pg.width[0];
        const double ci = floor(inbox / pg.h[0]);
        const double xr = (inbox - ci * pg.h[0]) / pg.h[0];
        const int jj = jx < W ? jx : 0;
        double a = 0.0;
        for (int q = deg; q >= 0; q--) a = a * xr + tab[jj * (deg + 1) + q];
        bxl = jx < W ? a : 0.0;
        int v = cell_of(s, pg.h[0], n) - NB + jj;
        v = v < 0 ? v + n : (v >= n ? v - n : v);
        ixl = v;
    }
    // real planes 2 plane_size(n) reals apart, y rows 2 row_stride(n) apart
    const size_t ps = 2 * plane_size(n), ys = 2 * row_stride(n);
#pragma unroll
    for (int c = 0; c < NC; c++) acc[c] = 0.0;
    if (jx < W) {
        for (int jz = 0; jz < W; jz++) {
            double col[NC];
#pragma unroll
            for (int c = 0; c < NC; c++) col[c] = 0.0;
#pragma unroll 2
            for (int jy = 0; jy < W; jy++) {
                const size_t pt = size_t(iy[jy]) * ys + size_t(iz[jz]) * n + ixl;
#pragma unroll
                for (int c = 0; c < NC; c++) col[c] = fma(by[jy], double(f[c * ps + pt]), col[c]);
            }
#pragma unroll
            for (int c = 0; c < NC; c++) acc[c] = fma(bz[jz], col[c], acc[c]);
        }
#pragma unroll
        for (int c = 0; c < NC; c++) acc[c] *= bxl;
    }
#pragma unroll
    for (int c = 0; c < NC; c++) {
#pragma unroll
        for (int o = LP / 2; o > 0; o /= 2) acc[c] += __shfl_xor(acc[c], o, LP);
    }
}

// One tracer of the Adams-Bashforth loop (MODE 0 of the brick kernel, for
// the global-gather path): velocity at the tracer into the history head,
// then the position update of order `order`.  t = rank of the tracer in the
// sorted order, jx = lane in its group.
template <typename T, int NB>
__device__ __forceinline__ void interp_ab_one(const T* __restrict__ f, double* __restrict__ pos,
                                              double* __restrict__ rhs, const uint32_t* __restrict__ perm,
                                              int64_t np, int nsteps, int head, int order, int add_head, double dt,
                                              const double* stab, int deg, const PartGeomD& pg, int64_t t, int jx) {
    const int64_t p = perm[t];
    double x[3] = {pos[p * 3 + 0], pos[p * 3 + 1], pos[p * 3 + 2]};
    double r0[3];
    gather_coop<T, NB, 3>(f, pg.n, x, stab, deg, pg, jx, r0);
    if (jx != 0) return;
    double* h0 = rhs + size_t(head) * np * 3 + p * 3;
    if (add_head) {
        r0[0] += h0[0];
        r0[1] += h0[1];
        r0[2] += h0[2];
    }
    h0[0] = r0[0];
    h0[1] = r0[1];
    h0[2] = r0[2];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double r[6];
        r[0] = r0[d];
        for (int i = 1; i < order; i++) r[i] = rhs[size_t((head + i) % nsteps) * np * 3 + p * 3 + d];
        double incr;
        switch (order) {
        case 1: incr = dt * r[0]; break;
        case 2: incr = dt * (3. * r[0] - r[1]) / 2.; break;
        case 3: incr = dt * (23. * r[0] - 16. * r[1] + 5. * r[2]) / 12.; break;
        case 4: incr = dt * (55. * r[0] - 59. * r[1] + 37. * r[2] - 9. * r[3]) / 24.; break;
        case 5: incr = dt * (1901. * r[0] - 2774. * r[1] + 2616. * r[2] - 1274. * r[3] + 251. * r[4]) / 720.; break;
        default: incr = dt * (4277. * r[0] - 7923. * r[1] + 9982. * r[2] - 7298. * r[3] + 2877. * r[4] - 475. * r[5]) / 1440.; break;
        }
        x[d] += incr;
    }
    pos[p * 3 + 0] = x[0];
    pos[p * 3 + 1] = x[1];
    pos[p * 3 + 2] = x[2];
}

template <typename T, int NB>
__global__ void __launch_bounds__(256)
k_interp_ab(const T* __restrict__ f, double* __restrict__ pos, double* __restrict__ rhs,
            const uint32_t* __restrict__ perm, int64_t np, int nsteps, int head, int order, int add_head,
            double dt, const double* __restrict__ tab, int deg, PartGeomD pg) {
    constexpr int LP = Lanes<NB>::LP;
    extern __shared__ double stab[];
    const int tabsize = (2 * NB + 2) * (deg + 1);  // derivative-0 block only
    for (int i = threadIdx.x; i < tabsize; i += blockDim.x) stab[i] = tab[i];
    __syncthreads();
    const int64_t t = (xcd_block() * blockDim.x + threadIdx.x) / LP;
    const int jx = threadIdx.x % LP;
    if (t >= np) return;  // whole lane groups exit together
    interp_ab_one<T, NB>(f, pos, rhs, perm, np, nsteps, head, order, add_head, dt, stab, deg, pg, t, jx);
}

// The tracer loops of many boxes (an ensemble) in one launch: box b owns
// workgroups [block0_b, block0_b + ceil(np_b LP / 256)) and its own field,
// tracers, geometry and time step (TracerBox, rewritten by the host only
// when one of them changes), and its Adams-Bashforth state (TracerDyn),
// which lives on the device: the loop reads it and k_tracer_dyn_advance
// moves it to the next step, so that a captured graph replays the loop
// step after step with no host-side arguments.
struct TracerBox {
    const void* f;
    double* pos;
    double* rhs;
    const uint32_t* perm;
    const double* tab;
    int64_t np;
    PartGeomD pg;
    double dt;
    int nsteps, deg;
    uint32_t block0;
};
struct TracerDyn {
    int step_idx;  // the order of the step is min(step_idx, nsteps)
    int head;      // history slot of the newest right-hand side
    int add_head;  // 1: add into the head slot (it holds this step's partial sum)
    int pad;
};
static_assert(sizeof(TracerDyn) == 4 * sizeof(int), "TracerDyn is mirrored on the host as 4 ints");

// inc_step_idx + shift_rhs_vectors (particles_system.hpp:191-208) of every
// box of a group, on the device
__global__ void k_tracer_dyn_advance(const TracerBox* __restrict__ boxes, TracerDyn* __restrict__ dyn, int nboxes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nboxes) return;
    TracerDyn d = dyn[i];
    const int ns = boxes[i].nsteps;
    d.step_idx += 1;
    d.head = (d.head + ns - 1) % ns;
    d.add_head = 0;
    dyn[i] = d;
}

template <typename T, int NB>
__global__ void __launch_bounds__(256)
k_interp_ab_multi(const TracerBox* __restrict__ boxes, const TracerDyn* __restrict__ dyn, int nboxes) {
    constexpr int LP = Lanes<NB>::LP;
    extern __shared__ double stab[];
    // the box of this workgroup: the last one starting at or before it
    int lo = 0, hi = nboxes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (boxes[mid].block0 <= blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const TracerBox& bx = boxes[lo];
    const int deg = bx.deg;
    const int tabsize = (2 * NB + 2) * (deg + 1);
    for (int i = threadIdx.x; i < tabsize; i += blockDim.x) stab[i] = bx.tab[i];
    __syncthreads();
    const int64_t t = (int64_t(blockIdx.x - bx.block0) * blockDim.x + threadIdx.x) / LP;
    const int jx = threadIdx.x % LP;
    if (t >= bx.np) return;  // whole lane groups exit together
    const PartGeomD pg = bx.pg;
    const TracerDyn d = dyn[lo];
    interp_ab_one<T, NB>((const T*)bx.f, bx.pos, bx.rhs, bx.perm, bx.np, bx.nsteps, d.head,
                         min(d.step_idx, bx.nsteps), d.add_head, bx.dt, stab, deg, pg, t, jx);
}

template <typename T, int NB, int NC>
__global__ void __launch_bounds__(256)
k_sample(const T* __restrict__ f, const double* __restrict__ pos, double* __restrict__ out,
         const uint32_t* __restrict__ perm, int64_t np, const double* __restrict__ tab, int deg, PartGeomD pg) {
    constexpr int LP = Lanes<NB>::LP;
    extern __shared__ double stab[];
    const int tabsize = (2 * NB + 2) * (deg + 1);
    for (int i = threadIdx.x; i < tabsize; i += blockDim.x) stab[i] = tab[i];
    __syncthreads();
    const int64_t t = (xcd_block() * blockDim.x + threadIdx.x) / LP;
    const int jx = threadIdx.x % LP;
    if (t >= np) return;
    const int64_t p = perm[t];
    const double x[3] = {pos[p * 3 + 0], pos[p * 3 + 1], pos[p * 3 + 2]};
    double acc[NC];
    gather_coop<T, NB, NC>(f, pg.n, x, stab, deg, pg, jx, acc);
    if (jx != 0) return;
#pragma unroll
    for (int c = 0; c < NC; c++) out[p * NC + c] = acc[c];
}

// ---------------------------------------------------------------- LDS bricks
// Tracers are bucketed by brick (KX x KY x KZ cells).  One workgroup per CU
// walks a fixed list of bricks; for each brick it stages the brick plus the
// stencil halo (n cells below, n+1 above, per axis, periodic) of one field
// component at a time in LDS, and groups of LPT = 16 RY lanes accumulate one
// tracer each from LDS: lane (jx, ry) holds stencil column x = jx and reads
// rows y = ry, ry + RY, ... of each z plane.  The tile is double-buffered
// and filled by LDS DMA (global_load_lds: no registers, all of a wave's loads
// in flight at once): while the groups sum component c from one buffer, the
// next component -- or the next brick's first -- arrives in the other.
//
// Interpolation weights: with STORE_W the batch prologue evaluates the
// 3 (2n+2) weights of each tracer once (Horner in FP64, one weight per
// thread) into LDS, and every component's sum reads them from there
// (broadcast reads); otherwise (when the tiles leave no room) the lanes
// evaluate them per component.  Either way the polynomial degree is a
// uniform branch, so no lane evaluates the unused high-order terms.
#ifndef BFPSH_BRICK_NT
#define BFPSH_BRICK_NT 512
#endif
#ifndef BFPSH_BRICK_RY
#define BFPSH_BRICK_RY 2
#endif
#ifndef BFPSH_BRICK_PAD
#define BFPSH_BRICK_PAD 1
#endif
template <typename T, int NB>
struct BrickCfg {
    static constexpr int W = 2 * NB + 2;
    static constexpr int DW = int(sizeof(T) / 4);  // dwords per value
    static constexpr int RY = BFPSH_BRICK_RY;      // y rows per LDS read
    static constexpr int LPT = 16 * RY;            // lanes per tracer
    static constexpr int NI = W / RY;              // y rows per lane
    static_assert(W % RY == 0, "stencil rows split evenly over the lane rows");
    // FP32: 16 x 8 x 8 bricks (16 planes deep for stencils up to 10 wide),
    // which leave room for the stored weights of a 64-tracer batch beside
    // two tiles; FP64: 8 x 8 x 8 bricks, weights per component
    static constexpr int KX = sizeof(T) == 4 ? 16 : 8;
    static constexpr int KY = 8;
    static constexpr int KZ = (sizeof(T) == 4 && W <= 10) ? 16 : 8;
    static constexpr int WX = KX + W - 1, WY = KY + W - 1, WZ = KZ + W - 1;
    // the tile's x rows are windows of whole 16-byte groups, aligned in
    // global memory (XA values each), so that one LDS-DMA lane moves a
    // group: the window starts XS values before the stencil's first column
    // and is WXT values wide
    static constexpr int XA = int(16 / sizeof(T));
    static constexpr int XS = (XA - NB % XA) % XA;
    static constexpr int WXT = (WX + XS + XA - 1) / XA * XA;
    // y-plane stride (values): with RY >= 2 the rows y and y+1 of a stencil
    // start 16 dwords apart modulo 32 banks, so the two 16-lane rows of a
    // read touch disjoint banks whatever the tracer's x offset
    static constexpr int ys_of(int v) {
        if (!BFPSH_BRICK_PAD || RY < 2) return v;
        int s = v;
        while ((s * DW) % 32 != 16) s++;
        return s;
    }
    static constexpr int YS = ys_of(WZ * WXT);
    static_assert(YS % XA == 0, "y planes start on 16-byte groups");
    static constexpr int TILE = WY * YS;                           // values, pads included
    static constexpr int CHUNKS = (TILE + 64 * XA - 1) / (64 * XA);  // 1 KB DMA chunks (one per wave load)
    static constexpr size_t TILE_BYTES = size_t(CHUNKS) * 1024;
    static constexpr int NT = BFPSH_BRICK_NT;
    static constexpr int KPW = (CHUNKS + NT / 64 - 1) / (NT / 64);  // chunks per wave
    static constexpr int GROUPS = NT / LPT;
    // tracers per batch: a brick's tracers are normally one batch (512^3,
    // 2^22 tracers: 32 per 16x8x8 brick on average); every batch refills
    // the tiles of all components
    static constexpr int PB = 64;
    static_assert(PB <= NT && NT % PB == 0, "whole groups of threads per tracer in the batch prologue");
    // stored weights: per tracer 3 axes x WS doubles, tracers WQ doubles
    // apart (WQ = 2 mod 16: the two tracers of a wave read other banks)
    static constexpr int WS = 16;
    static constexpr int WQ = 3 * WS + 2;
    // the derivative-0 weight table (W polynomials of degree <= max(2n+1, 5))
    static constexpr int TABD = W * ((W - 1 > 5 ? W - 1 : 5) + 1);
    template <int NC, int MODE> static constexpr size_t res_doubles() {
        return MODE == 0 ? size_t(PB) * NC : (size_t(PB) + 1) / 2;
    }
    // two tiles, then results [PB][NC] (or slots [PB], uint32) + in-cell
    // positions [PB][3] (doubles) + cells [PB][3] (ints) [+ weights [PB][WQ]]
    template <int NC, int MODE, bool SW> static constexpr size_t lds_bytes() {
        return 2 * TILE_BYTES + (res_doubles<NC, MODE>() + size_t(PB) * 3 + (SW ? size_t(PB) * WQ + TABD : 0)) * 8 +
               size_t(PB) * 3 * 4;
    }
    static constexpr size_t LDS_MAX = 160 * 1024;
    static constexpr bool STORE_W = lds_bytes<3, 0, true>() <= LDS_MAX && lds_bytes<9, 1, true>() <= LDS_MAX;
#ifdef BFPSH_NO_BRICKS
    static constexpr bool ok = false;
#else
    static constexpr bool ok = W <= 16 && lds_bytes<3, 0, STORE_W>() <= LDS_MAX && lds_bytes<9, 1, STORE_W>() <= LDS_MAX;
#endif
};

// Visiting order of the bricks.  Workgroup g takes logical bricks
// k G + rank(g), k = 0, 1, ..., with rank XCD-aware (blocks are dealt
// round-robin to the 8 XCDs, so the G/8 ranks of one XCD are consecutive).
// With G = 256 (one workgroup per CU of an MI355X) the 32 bricks one XCD
// takes in round k form a superblock of 4 x 4 x 2 bricks, and successive
// rounds of an XCD walk a column of superblocks along z: the halos the 32
// workgroups of an XCD share at a time (and the planes shared with the
// next superblock of the column) are read from HBM once and then from
// that XCD's L2.  Other grids fall back to groups of 2^3 bricks.
// Brick counts stay below 2^20 (1024^3 in 8^3 bricks), so the walk uses
// 32-bit arithmetic.
#ifndef BFPSH_BRICK_GROUP
#define BFPSH_BRICK_GROUP 2
#endif
// planes of LDS reads in flight ahead of the stencil sums
#ifndef BFPSH_BRICK_PREFETCH
#define BFPSH_BRICK_PREFETCH 2
#endif
__device__ __forceinline__ uint32_t brick_of(uint32_t logical, uint32_t nbx, uint32_t nby, uint32_t nbz,
                                             uint32_t G) {
    constexpr uint32_t SX = 4, SY = 4, SZ = 2, SB = SX * SY * SZ;
    if (G == 8 * SB && nbx % SX == 0 && nby % SY == 0 && nbz % SZ == 0 && ((nbx / SX) * (nby / SY)) % 8 == 0) {
        const uint32_t nsx = nbx / SX, nsz = nbz / SZ;
        const uint32_t k = logical / G, x = (logical % G) / SB, j = logical % SB;  // round, XCD, slot
        const uint32_t col = (k / nsz) * 8 + x, lev = k % nsz;
        const uint32_t bx = (col % nsx) * SX + j % SX;
        const uint32_t by = (col / nsx) * SY + (j / SX) % SY;
        const uint32_t bz = lev * SZ + j / (SX * SY);
        return (bz * nby + by) * nbx + bx;
    }
    constexpr uint32_t BG = BFPSH_BRICK_GROUP;
    if (BG <= 1 || nbx % BG || nby % BG || nbz % BG) return logical;
    constexpr uint32_t G3 = BG * BG * BG;
    const uint32_t g = logical / G3, w = logical % G3;
    const uint32_t ngx = nbx / BG, ngy = nby / BG;
    const uint32_t bx = (g % ngx) * BG + w % BG;
    const uint32_t by = ((g / ngx) % ngy) * BG + (w / BG) % BG;
    const uint32_t bz = (g / (ngx * ngy)) * BG + w / (BG * BG);
    return (bz * nby + by) * nbx + bx;
}

__global__ void k_brick_keys(const double* __restrict__ pos, int64_t np, PartGeomD pg, int kx, int ky, int kz,
                             int nbx, int nby, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int cx = cell_of(pos[p * 3 + 0], pg.h[0], pg.n) / kx;
    const int cy = cell_of(pos[p * 3 + 1], pg.h[1], pg.n) / ky;
    const int cz = cell_of(pos[p * 3 + 2], pg.h[2], pg.n) / kz;
    keys[p] = uint32_t((cz * nby + cy) * nbx + cx);
    vals[p] = uint32_t(p);
}

// [bstart, bend) of each brick in the sorted order (bricks without tracers
// keep the zero fill)
__global__ void k_brick_bounds(const uint32_t* __restrict__ skeys, int64_t np, uint32_t* __restrict__ bstart,
                               uint32_t* __restrict__ bend) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const uint32_t k = skeys[i];
    if (i == 0 || skeys[i - 1] != k) bstart[k] = uint32_t(i);
    if (i == np - 1 || skeys[i + 1] != k) bend[k] = uint32_t(i + 1);
}

// y weights of the lane rows and all z weights, spread within each 16-lane
// row by DPP broadcasts (lane k of a row holds the weight of index k)
template <int RY, int... I>
BFPSH_DEV void bcast_rows(double (&w)[sizeof...(I)], double v, std::integer_sequence<int, I...>) {
    ((w[I] = dpp_d<0x150 + I * RY>(v)), ...);
}

// Horner of degree D (compile time) on coefficients c[0..D], lowest first
template <int D, int DM>
BFPSH_DEV double horner(const double (&c)[DM + 1], double xr) {
    double a = c[D];
#pragma unroll
    for (int k = D - 1; k >= 0; k--) a = fma(a, xr, c[k]);
    return a;
}

// Polynomial weight of one stencil index at in-cell position xr; the
// lane's table row is in registers.  Degrees: 3 (m = 1), 5 (m = 2), 2n+1
// (m = 0, DM); the branch is uniform (deg is a kernel argument).
template <int DM>
BFPSH_DEV double weight_of(const double (&row)[DM + 1], int deg, double xr) {
    if (deg == 3) return horner<3, DM>(row, xr);
    if constexpr (DM >= 5) {
        if (deg == 5) return horner<5, DM>(row, xr);
    }
    return horner<DM, DM>(row, xr);
}

// weight j of one tracer on the three axes (in-cell positions fx, fy, fz)
// from the degree-D coefficients c of polynomial j, into w[0], w[WS], w[2 WS]
template <int D, int WS>
BFPSH_DEV void weights3(const double* __restrict__ c, double fx, double fy, double fz, double* __restrict__ w) {
    double k[D + 1];
#pragma unroll
    for (int i = 0; i <= D; i++) k[i] = c[i];
    double a = k[D], b = k[D], e = k[D];
#pragma unroll
    for (int i = D - 1; i >= 0; i--) {
        a = fma(a, fx, k[i]);
        b = fma(b, fy, k[i]);
        e = fma(e, fz, k[i]);
    }
    w[0] = a;
    w[WS] = b;
    w[2 * WS] = e;
}

// Adams-Bashforth move of tracer p from its position x by the new right-hand
// side r0 (already in the head slot of the history ring) and the order - 1
// older ones (particles_adams_bashforth.hpp:18-133)
BFPSH_DEV void ab_move(double* __restrict__ pos, const double* __restrict__ rhs, int64_t p, const double (&r0)[3],
                       const double (&x)[3], int64_t np, int nsteps, int head, int order, double dt) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double r[6];
        r[0] = r0[d];
        for (int i = 1; i < order; i++) r[i] = rhs[size_t((head + i) % nsteps) * np * 3 + p * 3 + d];
        double incr;
        switch (order) {
        case 1: incr = dt * r[0]; break;
        case 2: incr = dt * (3. * r[0] - r[1]) / 2.; break;
        case 3: incr = dt * (23. * r[0] - 16. * r[1] + 5. * r[2]) / 12.; break;
        case 4: incr = dt * (55. * r[0] - 59. * r[1] + 37. * r[2] - 9. * r[3]) / 24.; break;
        case 5: incr = dt * (1901. * r[0] - 2774. * r[1] + 2616. * r[2] - 1274. * r[3] + 251. * r[4]) / 720.; break;
        default: incr = dt * (4277. * r[0] - 7923. * r[1] + 9982. * r[2] - 7298. * r[3] + 2877. * r[4] - 475. * r[5]) / 1440.; break;
        }
        pos[p * 3 + d] = x[d] + incr;
    }
}

// one weight from the table in global memory (runtime degree; the batch
// prologue evaluates each weight once)
BFPSH_DEV double weight_rt(const double* __restrict__ tab, int j, int deg, double xr) {
    const double* c = tab + j * (deg + 1);
    double a = c[deg];
    for (int k = deg - 1; k >= 0; k--) a = fma(a, xr, c[k]);
    return a;
}

// MODE 0: velocity + Adams-Bashforth into the history ring; MODE 1: write
// the NC sampled components to out[p][NC].  One workgroup per CU: workgroup
// g takes the logical bricks k G + rank(g), k = 0, 1, ... (rank: XCD-aware,
// so that one XCD's workgroups hold neighbouring bricks at a time).  Work is
// a sequence of phases (brick, batch of up to PB tracers, component); the
// tile of phase i + 1 is requested before phase i is summed.  Per batch a
// prologue loads each tracer's position once (thread q owns tracer q),
// leaves its cell offset in the brick and in-cell fractions in LDS, and
// (STORE_W) evaluates its weights into LDS.  All control flow that reaches
// a barrier is uniform over the workgroup.
// Arithmetic as the reference's (particles_field_computer.hpp:142-150):
// FP64 weights, each FP32 or FP64 field value converted to double and
// accumulated in double.
template <typename T, int NB, int NC, int MODE>
__global__ void __launch_bounds__((BrickCfg<T, NB>::NT))
k_brick(const T* __restrict__ f, double* __restrict__ pos, double* __restrict__ rhs, double* __restrict__ out,
        const uint32_t* __restrict__ perm, const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ bend,
        int nbx, int nby, int64_t np, int nsteps, int head, int order, int add_head, double dt,
        const double* __restrict__ tab, int deg, PartGeomD pg) {
    using BC = BrickCfg<T, NB>;
    constexpr int W = BC::W, NT = BC::NT, PB = BC::PB, LPT = BC::LPT, RY = BC::RY, NI = BC::NI;
    constexpr int GROUPS = BC::GROUPS, WX = BC::WX, WY = BC::WY, WZ = BC::WZ, YS = BC::YS, DW = BC::DW;
    constexpr int KX = BC::KX, KY = BC::KY, KZ = BC::KZ, WS = BC::WS, WQ = BC::WQ;
    constexpr int XA = BC::XA, XS = BC::XS, WXT = BC::WXT, KPW = BC::KPW;
    constexpr bool SW = BC::STORE_W;
    // highest polynomial degree of an (n, m) table: 2n+1 (m = 0) or 5 (m = 2)
    constexpr int DM = W - 1 > 5 ? W - 1 : 5;
    extern __shared__ double smem[];
    unsigned char* tiles = reinterpret_cast<unsigned char*>(smem);  // 2 x TILE_BYTES
    double* res = smem + 2 * BC::TILE_BYTES / 8;                    // MODE 0: [PB][NC]
    uint32_t* lslot = reinterpret_cast<uint32_t*>(res);             // MODE 1: [PB]
    double* frac = res + BC::template res_doubles<NC, MODE>();      // [PB][3]
    double* wgt = frac + PB * 3;                                    // SW: [PB][WQ]
    double* tabl = wgt + (SW ? PB * WQ : 0);                        // SW: the weight table
    int* lcell = reinterpret_cast<int*>(tabl + (SW ? BC::TABD : 0)); // [PB][3]

    const int n = pg.n;
    const int nbz = n / KZ;
    const uint32_t nbricks = uint32_t(nbx) * uint32_t(nby) * uint32_t(nbz);
    const int G = gridDim.x;
    const int rank = (G % 8 == 0) ? int(blockIdx.x % 8) * (G / 8) + int(blockIdx.x / 8) : int(blockIdx.x);
    const int tid = threadIdx.x;
    const int lane = tid % LPT, grp = tid / LPT, jx = lane & 15, ry = lane >> 4;
    // real planes 2 plane_size(n) reals apart, y rows ys reals apart (the
    // offsets within a plane fit 32 bits for n <= 1024)
    const size_t ps = 2 * plane_size(n);
    const uint32_t ys = uint32_t(2 * row_stride(n));
    const uint32_t nxy = uint32_t(nbx) * uint32_t(nby);

    // without stored weights: this lane's row of the weight table (index
    // jx), in registers
    double trow[DM + 1];
    if constexpr (!SW) {
        const int jj = jx < W ? jx : 0;
#pragma unroll
        for (int k = 0; k <= DM; k++) trow[k] = k <= deg ? tab[jj * (deg + 1) + k] : 0.0;
    }

    // next non-empty brick at or after logical round kr (uniform)
    uint32_t kr = 0, brick = 0;
    uint32_t s0 = 0, e0 = 0;
    auto find = [&](uint32_t from, uint32_t& kout, uint32_t& bout, uint32_t& sout, uint32_t& eout) -> bool {
        for (uint32_t k = from;; k++) {
            const uint32_t L = k * uint32_t(G) + uint32_t(rank);
            if (L >= nbricks) return false;
            const uint32_t bb = brick_of(L, uint32_t(nbx), uint32_t(nby), uint32_t(nbz), uint32_t(G));
            const uint32_t a = bstart[bb], e = bend[bb];
            if (a < e) {
                kout = k;
                bout = bb;
                sout = a;
                eout = e;
                return true;
            }
        }
    };
    // request the tile of component c of brick bb into buffer b: LDS DMA of
    // one 16-byte group (XA values) per lane, 1 KB per wave load.  Lane ln
    // of this wave's k-th chunk holds the group at element offset loff[k]
    // from the window's origin (~0u: a pad slot, loaded from the origin);
    // bricks whose window crosses the box edge wrap each group instead.
    const int wv = tid / 64, ln = tid % 64;
    uint32_t loff[KPW];
#pragma unroll
    for (int k = 0; k < KPW; k++) {
        const int ch = wv + k * (NT / 64);
        const int v0 = (ch * 64 + ln) * XA;
        const int ly = v0 / YS, r = v0 % YS, lz = r / WXT, lx = r % WXT;
        loff[k] = (ch < BC::CHUNKS && ly < WY && lz < WZ) ? uint32_t(ly) * ys + uint32_t(lz * n + lx) : ~0u;
    }
    auto fill = [&](uint32_t bb, int c, int b) {
        const int bz = int(bb / nxy), by = int((bb % nxy) / uint32_t(nbx)), bx = int(bb % uint32_t(nbx));
        const int x0 = bx * KX - NB - XS, y0 = by * KY - NB, z0 = bz * KZ - NB;
        const T* fc = f + c * ps;
        uint32_t* tdw = reinterpret_cast<uint32_t*>(tiles + size_t(b) * BC::TILE_BYTES);
        const bool inside = x0 >= 0 && x0 + WXT <= n && y0 >= 0 && y0 + WY <= n && z0 >= 0 && z0 + WZ <= n;
        if (inside) {
            const T* org = fc + (uint32_t(y0) * ys + uint32_t(z0 * n + x0));
#pragma unroll
            for (int k = 0; k < KPW; k++) {
                const int ch = wv + k * (NT / 64);
                if (k + 1 < KPW || ch < BC::CHUNKS) {  // uniform per wave
                    const T* src = org + (loff[k] == ~0u ? 0u : loff[k]);
                    __builtin_amdgcn_global_load_lds(reinterpret_cast<const void*>(src),
                                                     (__attribute__((address_space(3))) void*)(tdw + ch * 256), 16,
                                                     0, 0);
                }
            }
        } else {
            for (int ch = wv; ch < BC::CHUNKS; ch += NT / 64) {
                const int v0 = (ch * 64 + ln) * XA;
                const int ly = v0 / YS, r = v0 % YS, lz = r / WXT, lx = r % WXT;
                const T* src = fc;
                if (ly < WY && lz < WZ) {
                    int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;  // gx: a group start, never inside a wrap
                    gx = gx < 0 ? gx + n : (gx >= n ? gx - n : gx);
                    gy = gy < 0 ? gy + n : (gy >= n ? gy - n : gy);
                    gz = gz < 0 ? gz + n : (gz >= n ? gz - n : gz);
                    src = fc + (uint32_t(gy) * ys + uint32_t(gz * n + gx));
                }
                __builtin_amdgcn_global_load_lds(reinterpret_cast<const void*>(src),
                                                 (__attribute__((address_space(3))) void*)(tdw + ch * 256), 16, 0, 0);
            }
        }
    };

    // batches: the current one (cb) and the one after it (nb), each a brick
    // and a range [bs, min(bs + PB, e0)) of its sorted tracers
    struct Bat {
        uint32_t kr, brick, s0, e0, bs;
    };
    auto advance = [&](const Bat& c, Bat& out) -> bool {
        out = c;
        out.bs = c.bs + PB;
        if (out.bs < c.e0) return true;  // more tracers of the same brick
        if (!find(c.kr + 1, out.kr, out.brick, out.s0, out.e0)) return false;
        out.bs = out.s0;
        return true;
    };
    if (!find(0, kr, brick, s0, e0)) return;  // uniform
    Bat cb{kr, brick, s0, e0, s0}, nb = cb;
    bool has_next = advance(cb, nb);
    // each thread stages one tracer of a batch ahead of time: tracer
    // tid % PB (every one of the NT / PB threads of a tracer evaluates a
    // share of its weights); the loads for batch b + 1 are issued during
    // batch b's component phases, so their latency hides behind its sums
    uint32_t pre_p = 0;
    double pre_x[3] = {0.0, 0.0, 0.0};
    {
        const int q = tid % PB, c0 = int(min(uint32_t(PB), cb.e0 - cb.bs));
        if (q < c0) {
            pre_p = perm[cb.bs + q];
#pragma unroll
            for (int d = 0; d < 3; d++) pre_x[d] = pos[size_t(pre_p) * 3 + d];
        }
    }
    if constexpr (SW) {
        // the first batch's prologue evaluates weights from the table before
        // the loop's first barrier: the copy must be whole before it starts
        for (int i = tid; i < W * (deg + 1); i += NT) tabl[i] = tab[i];
        __syncthreads();
    }
    int comp = 0, buf = 0;
    fill(cb.brick, 0, 0);
    int64_t myp = -1;
    double myx[3] = {0.0, 0.0, 0.0};
    for (;;) {
        const int cnt = int(min(uint32_t(PB), cb.e0 - cb.bs));
        if (comp == 0) {
            // batch prologue from the staged positions (the previous batch
            // ended with a barrier): cell offsets in the brick and in-cell
            // fractions (get_norm_pos_in_cell / pbc_field_layer), and with
            // STORE_W this thread's share of the tracer's weights
            const int base[3] = {int(cb.brick % uint32_t(nbx)) * KX, int((cb.brick % nxy) / uint32_t(nbx)) * KY,
                                 int(cb.brick / nxy) * KZ};
            const int q = tid % PB, part = tid / PB;
            myp = -1;
            if (q < cnt) {
                double fr[3];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const double sp = pre_x[d];
                    const double rep = floor(sp / pg.width[d]);
                    const double inbox = sp - rep * pg.width[d];
                    const double ci = floor(inbox / pg.h[d]);
                    fr[d] = (inbox - ci * pg.h[d]) / pg.h[d];
                }
                if (part == 0) {
                    myp = pre_p;
                    if constexpr (MODE == 1) lslot[q] = pre_p;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        myx[d] = pre_x[d];
                        frac[q * 3 + d] = fr[d];
                        lcell[q * 3 + d] = cell_of(pre_x[d], pg.h[d], n) - base[d];
                    }
                }
                if constexpr (SW) {
                    for (int j = part; j < W; j += NT / PB)
#pragma unroll
                        for (int a = 0; a < 3; a++) wgt[q * WQ + a * WS + j] = weight_rt(tabl, j, deg, fr[a]);
                }
            }
        }
        // the phase after this one
        Bat pbn = cb;
        int ncomp = comp + 1;
        bool more = true;
        if (ncomp == NC) {
            ncomp = 0;
            more = has_next;
            pbn = nb;
        }
        __builtin_amdgcn_s_waitcnt(0);  // this wave's DMA and staging loads have landed
        __syncthreads();                // the tile is whole; the other buffer is free
        if (has_next) {
            // stage the next batch: its slots in this batch's first phase,
            // its positions in the second (the first phase when NC = 1)
            const int q = tid % PB, ncnt = int(min(uint32_t(PB), nb.e0 - nb.bs));
            if (q < ncnt) {
                if (comp == 0) pre_p = perm[nb.bs + q];
                if (comp == (NC > 1 ? 1 : 0)) {
#pragma unroll
                    for (int d = 0; d < 3; d++) pre_x[d] = pos[size_t(pre_p) * 3 + d];
                }
            }
        }
        if (more) fill(pbn.brick, ncomp, buf ^ 1);
        const T* tile = reinterpret_cast<const T*>(tiles + size_t(buf) * BC::TILE_BYTES);
        for (int q = grp; q < cnt; q += GROUPS) {  // whole lane groups iterate together
            // FP64 weights of this lane: x index jx, y indices ry + i RY,
            // every z index
            double wxo, wyr[NI], wz[W];
            if constexpr (SW) {
                const double* wq = wgt + q * WQ;
                wxo = jx < W ? wq[jx] : 0.0;
#pragma unroll
                for (int i = 0; i < NI; i++) wyr[i] = wq[WS + ry + i * RY];
#pragma unroll
                for (int j = 0; j < W; j++) wz[j] = wq[2 * WS + j];
            } else {
                // lane jx evaluates index jx of each axis (the y weight of
                // its row ry: index jx + ry, by a row shift); the group
                // spreads them by DPP broadcasts
                double wyo = 0.0, wzo = 0.0;
                wxo = 0.0;
                if (jx < W) {
                    wxo = weight_of<DM>(trow, deg, frac[q * 3 + 0]);
                    wyo = weight_of<DM>(trow, deg, frac[q * 3 + 1]);
                    wzo = weight_of<DM>(trow, deg, frac[q * 3 + 2]);
                }
                double wyl = wyo;
                if constexpr (RY >= 2) {
                    const double s1 = dpp_d<0x101>(wyo);
                    wyl = ry == 1 ? s1 : wyl;
                }
                if constexpr (RY >= 4) {
                    const double s2 = dpp_d<0x102>(wyo), s3 = dpp_d<0x103>(wyo);
                    wyl = ry == 2 ? s2 : (ry == 3 ? s3 : wyl);
                }
                bcast_rows<RY>(wyr, wyl, std::make_integer_sequence<int, NI>{});
                bcast_rows<1>(wz, wzo, std::make_integer_sequence<int, W>{});
            }
            const int lc0 = lcell[q * 3 + 0], lc1 = lcell[q * 3 + 1], lc2 = lcell[q * 3 + 2];
            // this lane's column: x = lc0 + jx (lanes past the stencil read
            // column W - 1 with weight 0), rows lc1 + ry + i RY, plane lc2 + jz
            const T* tcol = tile + (XS + lc0 + (jx < W ? jx : W - 1)) + lc2 * WXT + (lc1 + ry) * YS;
            double acc = 0.0;
            // software pipeline: the LDS reads of plane jz + PD are issued
            // before plane jz's conversions and FMAs (two waves per SIMD
            // cannot hide the LDS latency by themselves); the loop is fully
            // unrolled, so tb lives in registers, PD + 1 planes at a time
            constexpr int PD = BFPSH_BRICK_PREFETCH;
            T tb[W][NI];
#pragma unroll
            for (int jz = 0; jz < PD && jz < W; jz++)
#pragma unroll
                for (int i = 0; i < NI; i++) tb[jz][i] = tcol[jz * WXT + i * RY * YS];
#pragma unroll
            for (int jz = 0; jz < W; jz++) {
                if (jz + PD < W) {
#pragma unroll
                    for (int i = 0; i < NI; i++) tb[jz + PD][i] = tcol[(jz + PD) * WXT + i * RY * YS];
                }
                const T* t = tb[jz];
                double c0 = 0.0, c1 = 0.0;
#pragma unroll
                for (int i = 0; i < NI; i += 2) {
                    c0 = fma(wyr[i], double(t[i]), c0);
                    if (i + 1 < NI) c1 = fma(wyr[i + 1], double(t[i + 1]), c1);
                }
                acc = fma(wz[jz], c0 + c1, acc);
            }
            acc = row16_sum(acc * wxo);
#pragma unroll
            for (int o = 16; o < LPT; o *= 2) acc += __shfl_xor(acc, o, LPT);
            if (lane == 0) {
                if constexpr (MODE == 1)
                    out[size_t(lslot[q]) * NC + comp] = acc;
                else
                    res[q * NC + comp] = acc;
            }
        }

        if (comp == NC - 1) {
            __syncthreads();  // the batch's results are complete; its LDS slots may be reused
            if constexpr (MODE == 0) {
                static_assert(NC == 3, "the Adams-Bashforth epilogue moves 3 coordinates");
                if (tid < cnt) {
                    const int q = tid;
                    const int64_t p = myp;
                    double r0[3] = {res[q * NC + 0], res[q * NC + 1], res[q * NC + 2]};
                    double* h0 = rhs + size_t(head) * np * 3 + p * 3;
                    if (add_head) {
                        r0[0] += h0[0];
                        r0[1] += h0[1];
                        r0[2] += h0[2];
                    }
                    h0[0] = r0[0];
                    h0[1] = r0[1];
                    h0[2] = r0[2];
                    ab_move(pos, rhs, p, r0, myx, np, nsteps, head, order, dt);
                }
            }
        }
        if (!more) break;
        if (ncomp == 0) {
            cb = nb;
            has_next = advance(cb, nb);
        }
        comp = ncomp;
        buf ^= 1;
    }
}

// ---------------------------------------------------------------------------
// Column walk: the tracer kernel for boxes of 64^3 and more.  A workgroup
// owns a column of KX x KZ cells (x, z) and walks it along y in steps of KY
// cells, one velocity component per walk.  The y planes of the stencil
// window (KY + W - 1 of them) sit in an LDS ring of R = 2 KY + W - 1 slots,
// so every plane of the column is moved from HBM once per walk while the
// KY planes of the next step arrive by LDS DMA during the current step's
// sums.  Compared with independent bricks (whose tiles carry the W - 1 halo
// planes of all three axes) this removes the y halo from the traffic:
// 512^3, n = 6: 5.6 KB per plane slot, (n + 13) planes per walk and column,
// 4.5 GB per launch against 22 GB for 16 x 8 x 8 bricks.
// Columns are split into y segments when there are fewer than two columns
// per CU.  Consecutive workgroups of an XCD take x-neighbour columns (x
// fastest) and walk y at about the same pace: a window row (KX + 2n + 1
// values) spans three 128-byte lines, its own and one of each x neighbour,
// so those lines are read from HBM once and then from that XCD's L2.
// Each slot is a z x x window [WZ][ZS] of the plane, ZS chosen so that the
// two 16-lane row phases of a tracer (NI rows apart) read disjoint banks.
#ifndef BFPSH_COL_NT
#define BFPSH_COL_NT 1024
#endif
template <typename T, int NB>
struct ColCfg {
    static constexpr int W = 2 * NB + 2;
    static constexpr int DW = int(sizeof(T) / 4);
    static constexpr int RZ = 2;         // 16-lane row phases per tracer
    static constexpr int LPT = 16 * RZ;  // lanes per tracer
    static constexpr int NI = W / RZ;    // z rows per lane: phase rz takes rows rz NI .. rz NI + NI - 1
    static constexpr int KX = 32, KZ = 16;
    static constexpr int WX = KX + W - 1, WZ = KZ + W - 1;
    static constexpr int XA = int(16 / sizeof(T));
    static constexpr int XS = (XA - NB % XA) % XA;
    static constexpr int WXT = (WX + XS + XA - 1) / XA * XA;
    static constexpr int zs_of(int v) {
        for (int s = v; s < v + 64 * XA; s += XA)
            if ((NI * s * DW) % 32 == 16) return s;
        return v;
    }
    static constexpr int ZS = zs_of(WXT);
    static constexpr int SLOT = WZ * ZS;                       // values per slot, pads included
    static constexpr int CH = (SLOT + 64 * XA - 1) / (64 * XA);  // 1 KB DMA chunks per slot
    static constexpr int SLOT_BYTES = CH * 1024;
    // 1024 threads: one workgroup holds a CU's LDS, and 16 waves (4 per
    // SIMD) hide the LDS latency of the sums better than 8
    static constexpr int NT = BFPSH_COL_NT, GROUPS = NT / LPT, PB = 64, WS = 16, WQ = 3 * WS + 2;
    static constexpr int TABD = W * ((W - 1 > 5 ? W - 1 : 5) + 1);
    // y steps per work item (n / KY for an unsplit column)
    static constexpr int MAXS = 256;
    // weights [PB][WQ], table, results [PB] (doubles), cells [PB][3] (ints),
    // step ranges [2][MAXS] (uint32), ahead of the ring: at LDS offsets an
    // instruction's offset field carries, and the ring 1 KB aligned
    static constexpr size_t EXTRA = (size_t(PB) * WQ + TABD + PB) * 8 + size_t(PB) * 3 * 4 + size_t(2) * MAXS * 4;
    static constexpr size_t HEAD = (EXTRA + 1023) / 1024 * 1024;
    static constexpr size_t LDS_MAX = 160 * 1024;
    static constexpr size_t lds_of(int ky) { return size_t(2 * ky + W - 1) * SLOT_BYTES + HEAD; }
    static constexpr int KY = lds_of(8) <= LDS_MAX ? 8 : (lds_of(4) <= LDS_MAX ? 4 : 0);
    static constexpr int R = 2 * KY + W - 1;
    static constexpr size_t LDS = lds_of(KY);
#ifdef BFPSH_NO_COLUMNS
    static constexpr bool ok = false;
#else
    static constexpr bool ok = KY >= 4 && W <= 16 && CH <= NT / 64 && W % RZ == 0;
#endif
};

// sort key of a tracer, (column, y step) with columns z-fastest, and its
// record for the walk: in-cell fractions (get_norm_pos_in_cell /
// pbc_field_layer, as k_brick's prologue) and the cell, 10 bits per axis
__global__ void k_col_keys(const double* __restrict__ pos, int64_t np, PartGeomD pg, int kx, int ky, int kz,
                           int ncx, int nys, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                           double* __restrict__ rec) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int cell[3];
    double fr[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double sp = pos[p * 3 + d];
        const double rep = floor(sp / pg.width[d]);
        const double inbox = sp - rep * pg.width[d];
        const double ci = floor(inbox / pg.h[d]);
        fr[d] = (inbox - ci * pg.h[d]) / pg.h[d];
        cell[d] = cell_of(sp, pg.h[d], pg.n);
    }
    keys[p] = uint32_t(((cell[2] / kz) * ncx + cell[0] / kx) * nys + cell[1] / ky);
    vals[p] = uint32_t(p);
    double4 r;
    r.x = fr[0];
    r.y = fr[1];
    r.z = fr[2];
    r.w = __longlong_as_double((long long)(uint32_t(cell[0]) | (uint32_t(cell[1]) << 10) | (uint32_t(cell[2]) << 20)));
    reinterpret_cast<double4*>(rec)[p] = r;
}

// The NC components of a field at every tracer (k_brick's work, same
// arithmetic: FP64 weights, each FP32 field value converted to double and
// accumulated in double, particles_field_computer.hpp:142-150).  MODE 0:
// the velocity into the head slot of the history ring (k_ab_moves then
// moves every tracer: a coalesced pass over the history, kept out of the
// walks so that no workgroup waits on its loads); MODE 1: out[p][NC].
// Work item wi = seg ncols + col (col = cz ncx + cx) for wi = rank,
// rank + G, ...; per work item NC walks (one per component) over the
// segment's y steps [s0, s1), whose tracer ranges are read into LDS once
// per work item.  Per batch, thread t holds tracer t mod 64: its slot and
// its record (in-cell fractions and cell, written by k_col_keys) are
// loaded one and two batches ahead, so a batch starts with no global
// latency; each thread evaluates one weight index of its tracer on the
// three axes and the first wave leaves the cell offsets in LDS.  All
// control flow that reaches a barrier is uniform over the workgroup.
template <typename T, int NB, int NC, int MODE>
__global__ void __launch_bounds__((ColCfg<T, NB>::NT))
k_column(const T* __restrict__ f, const double* __restrict__ crec, double* __restrict__ out,
         const uint32_t* __restrict__ perm, const uint32_t* __restrict__ kstart, const uint32_t* __restrict__ kend,
         int ncx, int nys, int ncols, int sps, int nwork, int64_t np, int head, int add_head,
         const double* __restrict__ tab, int deg, PartGeomD pg) {
    using CC = ColCfg<T, NB>;
    constexpr int W = CC::W, NT = CC::NT, PB = CC::PB, LPT = CC::LPT, RZ = CC::RZ, NI = CC::NI;
    constexpr int GROUPS = CC::GROUPS, WZ = CC::WZ, ZS = CC::ZS, XA = CC::XA, XS = CC::XS, WXT = CC::WXT;
    constexpr int KX = CC::KX, KY = CC::KY, KZ = CC::KZ, WS = CC::WS, WQ = CC::WQ, R = CC::R, CH = CC::CH;
    constexpr int MAXS = CC::MAXS;
    constexpr int SLOTV = CC::SLOT_BYTES / int(sizeof(T));
    // highest polynomial degree of an (n, m) table: 2n+1 (m = 0) or 5 (m = 2)
    constexpr int DM = W - 1 > 5 ? W - 1 : 5;
    extern __shared__ double smem[];
    double* wgt = smem;                                             // [PB][WQ]
    double* tabl = wgt + PB * WQ;                                   // the weight table
    double* res = tabl + CC::TABD;                                  // [PB]
    int* lcell = reinterpret_cast<int*>(res + PB);                  // [PB][3]: origin, planes to the ring end
    uint32_t* rng = reinterpret_cast<uint32_t*>(lcell + PB * 3);    // [2][MAXS]: step ranges of the work item
    unsigned char* ring = reinterpret_cast<unsigned char*>(smem) + CC::HEAD;  // R x SLOT_BYTES

    const int n = pg.n;
    const int G = gridDim.x;
    const int rank = (G % 8 == 0) ? int(blockIdx.x % 8) * (G / 8) + int(blockIdx.x / 8) : int(blockIdx.x);
    const int tid = threadIdx.x;
    const int lane = tid % LPT, grp = tid / LPT, jx = lane & 15, rz = lane >> 4;
    const int wv = tid / 64, ln = tid % 64;
    // real planes 2 plane_size(n) reals apart, y rows ys reals apart (the
    // offsets within a plane fit 32 bits for n <= 1024)
    const size_t ps = 2 * plane_size(n);
    const uint32_t ys = uint32_t(2 * row_stride(n));
    // the weight table: first read after the first walk's opening barrier
    for (int i = tid; i < W * (deg + 1); i += NT) tabl[i] = tab[i];

    // work item -> column origin and y steps
    struct Item {
        int cx0, cz0, s0, s1;
    };
    auto item = [&](int wi) -> Item {
        const int col = wi % ncols, seg = wi / ncols;
        const int s0 = seg * sps;
        return Item{(col % ncx) * KX, (col / ncx) * KZ, s0, min(s0 + sps, nys)};
    };
    auto key_of = [&](int wi, int s) -> uint32_t { return uint32_t(wi % ncols) * uint32_t(nys) + uint32_t(s); };
    // tracer range of step s of work item wi: from the LDS table when wi is
    // the current work item cur (uniform)
    auto range = [&](int wi, int cur, int s0, int s, uint32_t& a, uint32_t& e) {
        if (wi == cur) {
            a = rng[s - s0];
            e = rng[MAXS + s - s0];
        } else {
            const uint32_t k = key_of(wi, s);
            a = kstart[k];
            e = kend[k];
        }
    };

    // batch cursor over (work item, walk, step, batch start); uniform
    struct Cur {
        int wi, c, s, s0, s1;
        uint32_t bs, e;
    };
    // first non-empty step at or after cu.s in walks c.. of work items wi..
    auto settle = [&](Cur& cu, int cur) -> bool {
        for (;;) {
            if (cu.wi >= nwork) return false;
            for (; cu.s < cu.s1; cu.s++) {
                uint32_t a, e;
                range(cu.wi, cur, cu.s0, cu.s, a, e);
                if (a < e) {
                    cu.bs = a;
                    cu.e = e;
                    return true;
                }
            }
            if (++cu.c == NC) {
                cu.c = 0;
                cu.wi += G;
                if (cu.wi >= nwork) return false;
                const Item it = item(cu.wi);
                cu.s0 = it.s0;
                cu.s1 = it.s1;
            }
            cu.s = cu.s0;
        }
    };
    auto next = [&](Cur& cu, int cur) -> bool {
        cu.bs += PB;
        if (cu.bs < cu.e) return true;
        cu.s++;
        return settle(cu, cur);
    };

    // LDS DMA: wave wv < CH moves chunk wv of each plane slot; lane ln holds
    // the 16-byte group at offset loff from the plane's column origin
    // (~0u: a pad slot, loaded from the origin).  x and z wrap per group
    // (x groups never straddle the box edge: n and the window start are
    // multiples of XA).
    uint32_t loff = ~0u;
    auto set_column = [&](const Item& it) {
        const int v0 = (wv * 64 + ln) * XA;
        const int lz = v0 / ZS, lx = v0 % ZS;
        loff = ~0u;
        if (wv < CH && lz < WZ && lx < WXT) {
            int gx = it.cx0 - NB - XS + lx, gz = it.cz0 - NB + lz;
            gx = gx < 0 ? gx + n : (gx >= n ? gx - n : gx);
            gz = gz < 0 ? gz + n : (gz >= n ? gz - n : gz);
            loff = uint32_t(gz * n + gx);
        }
    };
    // plane P of the walk (global y = s0 KY - NB + P, wrapped) into slot P % R
    auto load_plane = [&](const T* fc, int ybase, int P) {
        if (wv < CH) {  // uniform per wave
            int gy = ybase + P;
            gy = gy < 0 ? gy + n : (gy >= n ? gy - n : gy);
            const T* src = fc + uint32_t(gy) * ys + (loff == ~0u ? 0u : loff);
            unsigned char* dst = ring + size_t(P % R) * CC::SLOT_BYTES + size_t(wv) * 1024;
            __builtin_amdgcn_global_load_lds(reinterpret_cast<const void*>(src),
                                             (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
        }
    };

    // staging: thread t holds tracer t mod PB of the next batch (slot p1,
    // record x1) and the slot p2 of the batch after it
    const int qq = tid % PB;
    Cur n1{rank, 0, 0, 0, 0, 0, 0};
    if (n1.wi < nwork) {
        const Item it = item(n1.wi);
        n1.s = n1.s0 = it.s0;
        n1.s1 = it.s1;
    }
    bool h1 = settle(n1, -1);
    Cur n2 = n1;
    bool h2 = h1 && next(n2, -1);
    uint32_t p1 = 0, p2 = 0;
    double4 x1 = make_double4(0.0, 0.0, 0.0, 0.0);
    const double4* rec4 = reinterpret_cast<const double4*>(crec);
    if (h1 && n1.bs + qq < n1.e) {
        p1 = perm[n1.bs + qq];
        x1 = rec4[p1];
    }
    if (h2 && n2.bs + qq < n2.e) p2 = perm[n2.bs + qq];

    for (int wi = rank; wi < nwork; wi += G) {
        const Item it = item(wi);
        set_column(it);
        const int ybase = it.s0 * KY - NB;
        for (int i = tid; i < it.s1 - it.s0; i += NT) {
            const uint32_t k = key_of(wi, it.s0 + i);
            rng[i] = kstart[k];
            rng[MAXS + i] = kend[k];
        }
        for (int c = 0; c < NC; c++) {
            const T* fc = f + c * ps;
            __syncthreads();  // the previous walk's sums are done with the ring; the range table is whole
            for (int P = 0; P < KY + W - 1; P++) load_plane(fc, ybase, P);
            for (int s = it.s0; s < it.s1; s++) {
                __builtin_amdgcn_s_waitcnt(0);  // this wave's DMA and staging loads have landed
                __syncthreads();                // step s's window is whole; step s - 1's first slots are free
                if (s + 1 < it.s1) {
                    const int P0 = (s + 1 - it.s0) * KY + W - 1;
                    for (int k = 0; k < KY; k++) load_plane(fc, ybase, P0 + k);
                }
                const uint32_t a = rng[s - it.s0], e = rng[MAXS + s - it.s0];
                for (uint32_t bs = a; bs < e; bs += PB) {
                    // this batch is n1 (uniform): take its staged tracers,
                    // then stage the next two
                    const int cnt = int(min(uint32_t(PB), e - bs));
                    const uint32_t myp = p1;
                    const double4 myr = x1;
                    n1 = n2;
                    h1 = h2;
                    p1 = p2;
                    h2 = h2 && next(n2, wi);
                    if (h1 && n1.bs + qq < n1.e) x1 = rec4[p1];
                    if (h2 && n2.bs + qq < n2.e) p2 = perm[n2.bs + qq];
                    {
                        // thread (q, j) evaluates weight j of tracer q on
                        // the three axes from its table row in registers;
                        // the first wave leaves the cell offsets in the
                        // column window
                        static_assert(NT / PB >= W, "one weight index per thread");
                        const int q = qq, j = tid / PB;
                        if (q < cnt && j == 0) {
                            // the tracer's stencil origin in the window of
                            // its first plane's ring slot, and the number of
                            // its planes before the ring's end
                            const uint32_t pc = uint32_t(__double_as_longlong(myr.w));
                            const int lc0 = int(pc & 1023u) - it.cx0;
                            const int lc1 = int((pc >> 10) & 1023u) - it.s0 * KY;
                            const int lc2 = int((pc >> 20) & 1023u) - it.cz0;
                            const int sl0 = lc1 % R;
                            lcell[q * 3 + 0] = XS + lc0 + lc2 * ZS + sl0 * SLOTV;
                            lcell[q * 3 + 1] = R - sl0;
                        }
                        if (q < cnt && j < W) {
                            const double* c = tabl + j * (deg + 1);
                            double* wo = wgt + q * WQ + j;
                            const double fx = myr.x, fy = myr.y, fz = myr.z;
                            if (deg == 5) {
                                weights3<5, WS>(c, fx, fy, fz, wo);
                            } else if (deg == 3) {
                                weights3<3, WS>(c, fx, fy, fz, wo);
                            } else {
                                double trow[DM + 1];
#pragma unroll
                                for (int k = 0; k <= DM; k++) trow[k] = k <= deg ? c[k] : 0.0;
                                wo[0] = weight_of<DM>(trow, deg, fx);
                                wo[WS] = weight_of<DM>(trow, deg, fy);
                                wo[2 * WS] = weight_of<DM>(trow, deg, fz);
                            }
                        }
                    }
                    __syncthreads();  // weights and cells of the batch
                    // this lane's column x = origin + jx (lanes past the
                    // stencil read column W - 1 with weight 0) and z rows
                    // origin + rz NI + i (consecutive rows: one base
                    // address serves ds_read2 pairs)
                    const T* rt = reinterpret_cast<const T*>(ring) + (jx < W ? jx : W - 1) + rz * NI * ZS;
                    for (int qb = grp; qb < cnt; qb += GROUPS) {  // whole lane groups iterate together
                        const double* wq = wgt + qb * WQ;
                        const double wxo = jx < W ? wq[jx] : 0.0;
                        double wzr[NI];
#pragma unroll
                        for (int i = 0; i < NI; i++) wzr[i] = wq[2 * WS + rz * NI + i];
                        // plane jy of the stencil sits in ring slot
                        // (sl0 + jy) mod R: from pa before the ring's end,
                        // from pb = pa - R slots after it
                        const T* pa = rt + lcell[qb * 3 + 0];
                        const T* pb = pa - R * SLOTV;
                        const int wrap = lcell[qb * 3 + 1];
                        double acc = 0.0;
                        // software pipeline: plane jy + PD's LDS reads are
                        // issued before plane jy's conversions and FMAs
                        constexpr int PD = BFPSH_BRICK_PREFETCH;
                        // (the plane's weight is read with its values: LDS
                        // reads return in order, so waiting for plane jy
                        // never waits for the planes behind it)
                        T tb[W][NI];
                        double wyv[W];
                        auto rd = [&](int jy) {
                            const T* tz = (jy < wrap ? pa : pb) + jy * SLOTV;
                            wyv[jy] = wq[WS + jy];
#pragma unroll
                            for (int i = 0; i < NI; i++) tb[jy][i] = tz[i * ZS];
                        };
#pragma unroll
                        for (int jy = 0; jy < PD && jy < W; jy++) rd(jy);
#pragma unroll
                        for (int jy = 0; jy < W; jy++) {
                            if (jy + PD < W) rd(jy + PD);
                            double cz = 0.0;
#pragma unroll
                            for (int i = 0; i < NI; i++) cz = fma(wzr[i], double(tb[jy][i]), cz);
                            acc = fma(wyv[jy], cz, acc);
                        }
                        acc = row16_sum(acc * wxo);
#pragma unroll
                        for (int o = 16; o < LPT; o *= 2) acc += __shfl_xor(acc, o, LPT);
                        if (lane == 0) res[qb] = acc;
                    }
                    __syncthreads();  // the batch's results are complete; its LDS slots may be reused
                    if (tid < cnt) {
                        if constexpr (MODE == 0) {
                            static_assert(NC == 3, "the velocity loop interpolates 3 components");
                            double* h0 = out + size_t(head) * np * 3 + int64_t(myp) * 3 + c;
                            *h0 = res[tid] + (add_head ? *h0 : 0.0);
                        } else {
                            out[int64_t(myp) * NC + c] = res[tid];
                        }
                    }
                }
            }
        }
    }
}

// Adams-Bashforth move of every tracer from the right-hand side in the head
// slot of the history ring (the second half of the column walk's velocity
// loop)
__global__ void k_ab_moves(double* __restrict__ pos, const double* __restrict__ rhs, int64_t np, int nsteps,
                           int head, int order, double dt) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const double* h0 = rhs + size_t(head) * np * 3 + p * 3;
    const double r0[3] = {h0[0], h0[1], h0[2]};
    const double x[3] = {pos[p * 3 + 0], pos[p * 3 + 1], pos[p * 3 + 2]};
    ab_move(pos, rhs, p, r0, x, np, nsteps, head, order, dt);
}

constexpr int kThreads = 256;

unsigned blocks_for(int64_t np, int lp) {
    int64_t nb = (np * lp + kThreads - 1) / kThreads;
    nb = (nb + 7) / 8 * 8;  // whole rounds over the 8 XCDs (xcd_block)
    return unsigned(nb);
}

int ensure_sort(bfpsh_ctx* ctx) {
    if (ctx->d_keys) return BFPSH_OK;
    const int64_t np = ctx->np;
    HIPCHK(hipMalloc((void**)&ctx->d_keys, 2 * np * sizeof(uint32_t)));
    HIPCHK(hipMalloc((void**)&ctx->d_perm, 2 * np * sizeof(uint32_t)));
    size_t tmp = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, ctx->d_keys, ctx->d_keys + np, ctx->d_perm,
                                              ctx->d_perm + np, int(np), 0, 30, ctx->stream));
    HIPCHK(hipMalloc(&ctx->d_sort_tmp, tmp));
    ctx->sort_tmp_bytes = tmp;
    ctx->bytes += 4 * np * sizeof(uint32_t) + tmp;
    ctx->sort_age = 1 << 30;
    return BFPSH_OK;
}

// sorted permutation in d_perm + np
int sort_tracers(bfpsh_ctx* ctx) {
    int r = ensure_sort(ctx);
    if (r) return r;
    const int64_t np = ctx->np;
    const int shift = ctx->n >= 64 ? 2 : 0;
    int nbits = 0;
    while ((1 << nbits) < (ctx->n >> shift)) nbits++;
    nbits *= 3;
    LAUNCH("part_keys", (k_part_keys<<<unsigned((np + 255) / 256), 256, 0, ctx->stream>>>(
                            ctx->d_pos, np, pgeom_d(ctx), shift, ctx->d_keys, ctx->d_perm)));
    size_t tmp = ctx->sort_tmp_bytes;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(ctx->d_sort_tmp, tmp, ctx->d_keys, ctx->d_keys + np, ctx->d_perm,
                                              ctx->d_perm + np, int(np), 0, std::max(nbits, 1), ctx->stream));
    ctx->sort_age = 0;
    return BFPSH_OK;
}

template <typename T, int NB>
int ensure_bricks(bfpsh_ctx* ctx, int64_t nb) {
    if (ctx->d_bstart && ctx->nbricks == nb) return BFPSH_OK;
    if (ctx->d_bstart) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipFree(ctx->d_bstart));
        ctx->d_bstart = nullptr;
    }
    HIPCHK(hipMalloc((void**)&ctx->d_bstart, 2 * nb * sizeof(uint32_t)));
    ctx->nbricks = nb;
    ctx->bytes += 2 * nb * sizeof(uint32_t);
    return BFPSH_OK;
}

// bucket the tracers by brick: sorted permutation in d_perm + np, per-brick
// ranges in d_bstart[0..nb), d_bstart[nb..2nb)
template <typename T, int NB>
int sort_bricks(bfpsh_ctx* ctx) {
    using BC = BrickCfg<T, NB>;
    int r = ensure_sort(ctx);
    if (r) return r;
    const int n = ctx->n;
    const int nbx = n / BC::KX, nby = n / BC::KY, nbz = n / BC::KZ;
    const int64_t nb = int64_t(nbx) * nby * nbz;
    r = ensure_bricks<T, NB>(ctx, nb);
    if (r) return r;
    const int64_t np = ctx->np;
    int nbits = 1;
    while ((int64_t(1) << nbits) < nb) nbits++;
    HIPCHK(hipMemsetAsync(ctx->d_bstart, 0, 2 * nb * sizeof(uint32_t), ctx->stream));
    LAUNCH("part_keys", (k_brick_keys<<<unsigned((np + 255) / 256), 256, 0, ctx->stream>>>(
                            ctx->d_pos, np, pgeom_d(ctx), BC::KX, BC::KY, BC::KZ, nbx, nby, ctx->d_keys, ctx->d_perm)));
    size_t tmp = ctx->sort_tmp_bytes;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(ctx->d_sort_tmp, tmp, ctx->d_keys, ctx->d_keys + np, ctx->d_perm,
                                              ctx->d_perm + np, int(np), 0, nbits, ctx->stream));
    LAUNCH("part_bounds", (k_brick_bounds<<<unsigned((np + 255) / 256), 256, 0, ctx->stream>>>(
                              ctx->d_keys + np, np, ctx->d_bstart, ctx->d_bstart + nb)));
    ctx->sort_age = 0;
    return BFPSH_OK;
}

template <typename T, int NB, int NC, int MODE>
int launch_brick(bfpsh_ctx* ctx, const void* f, double* d_out, double dt, int order, const char* name) {
    using BC = BrickCfg<T, NB>;
    const int n = ctx->n;
    const int nbx = n / BC::KX, nby = n / BC::KY;
    const int64_t nb = int64_t(nbx) * nby * (n / BC::KZ);
    const size_t lds = BC::template lds_bytes<NC, MODE, BC::STORE_W>();
    auto kern = k_brick<T, NB, NC, MODE>;
    HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    // one workgroup per CU (each takes the LDS of a CU), persistent over
    // its bricks
    int ncu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const unsigned grid = unsigned(std::max<int64_t>(1, std::min<int64_t>(nb, ncu)));
    LAUNCH(name, (kern<<<grid, BC::NT, lds, ctx->stream>>>(
                     (const T*)f, ctx->d_pos, ctx->d_rhs, d_out, ctx->d_perm + ctx->np, ctx->d_bstart,
                     ctx->d_bstart + nb, nbx, nby, ctx->np, ctx->nsteps, ctx->head, order, ctx->add_head ? 1 : 0, dt,
                     ctx->d_tab, ctx->deg, pgeom_d(ctx))));
    return BFPSH_OK;
}

// column walk geometry of a box: columns, y steps, steps per segment, work
// items (segments split the columns when there are fewer than two per CU)
struct ColGeom {
    int ncx, ncz, nys, ncols, sps, nwork;
};
template <typename T, int NB>
bool col_geom(const bfpsh_ctx* ctx, ColGeom& g) {
    using CC = ColCfg<T, NB>;
    static_assert(CC::ok, "column walk geometry of a configuration that has none");
    const int n = ctx->n;
    // (cells are packed 10 bits per axis)
    if (ctx->force_coop || (ctx->tracer_kernel != 0 && ctx->tracer_kernel != 3) || n < 64 || n > 1024 ||
        n % CC::KX || n % CC::KZ || n % CC::KY)
        return false;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) return false;
    g.ncx = n / CC::KX;
    g.ncz = n / CC::KZ;
    g.nys = n / CC::KY;
    g.ncols = g.ncx * g.ncz;
    const int want = (2 * ncu + g.ncols - 1) / g.ncols;  // segments per column
    const int nseg = std::max(1, std::min(g.nys, want));
    g.sps = (g.nys + nseg - 1) / nseg;
    g.nwork = g.ncols * ((g.nys + g.sps - 1) / g.sps);
    return g.sps <= CC::MAXS;
}

// bucket the tracers by (column, y step): sorted permutation in d_perm + np,
// per-key ranges in d_bstart[0..nk), d_bstart[nk..2nk)
template <typename T, int NB>
int sort_cols(bfpsh_ctx* ctx, const ColGeom& g) {
    using CC = ColCfg<T, NB>;
    int r = ensure_sort(ctx);
    if (r) return r;
    const int64_t nk = int64_t(g.ncols) * g.nys;
    r = ensure_bricks<T, NB>(ctx, nk);
    if (r) return r;
    const int64_t np = ctx->np;
    if (!ctx->d_crec) {
        HIPCHK(hipMalloc((void**)&ctx->d_crec, size_t(np) * 4 * sizeof(double)));
        ctx->bytes += size_t(np) * 4 * sizeof(double);
    }
    int nbits = 1;
    while ((int64_t(1) << nbits) < nk) nbits++;
    HIPCHK(hipMemsetAsync(ctx->d_bstart, 0, 2 * nk * sizeof(uint32_t), ctx->stream));
    LAUNCH("part_keys", (k_col_keys<<<unsigned((np + 255) / 256), 256, 0, ctx->stream>>>(
                            ctx->d_pos, np, pgeom_d(ctx), CC::KX, CC::KY, CC::KZ, g.ncx, g.nys, ctx->d_keys,
                            ctx->d_perm, ctx->d_crec)));
    size_t tmp = ctx->sort_tmp_bytes;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(ctx->d_sort_tmp, tmp, ctx->d_keys, ctx->d_keys + np, ctx->d_perm,
                                              ctx->d_perm + np, int(np), 0, nbits, ctx->stream));
    LAUNCH("part_bounds", (k_brick_bounds<<<unsigned((np + 255) / 256), 256, 0, ctx->stream>>>(
                              ctx->d_keys + np, np, ctx->d_bstart, ctx->d_bstart + nk)));
    ctx->sort_age = 0;
    return BFPSH_OK;
}

template <typename T, int NB, int NC, int MODE>
int launch_col(bfpsh_ctx* ctx, const ColGeom& g, const void* f, double* d_out, const char* name) {
    using CC = ColCfg<T, NB>;
    const int64_t nk = int64_t(g.ncols) * g.nys;
    auto kern = k_column<T, NB, NC, MODE>;
    HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(CC::LDS)));
    int ncu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    // one workgroup per CU, persistent over its work items
    const unsigned grid = unsigned(std::max(1, std::min(g.nwork, ncu)));
    LAUNCH(name, (kern<<<grid, CC::NT, CC::LDS, ctx->stream>>>(
                     (const T*)f, ctx->d_crec, d_out, ctx->d_perm + ctx->np, ctx->d_bstart, ctx->d_bstart + nk, g.ncx,
                     g.nys, g.ncols, g.sps, g.nwork, ctx->np, ctx->head, ctx->add_head ? 1 : 0, ctx->d_tab, ctx->deg,
                     pgeom_d(ctx))));
    return BFPSH_OK;
}

// Kernel forms of the velocity loop and the samples (BFPSH_TRACER_KERNEL
// forces one: brick, gather or column).
enum { FORM_AUTO = 0, FORM_BRICK = 1, FORM_GATHER = 2, FORM_COLUMN = 3 };

template <typename T, int NB>
bool form_ok(const bfpsh_ctx* ctx, int form) {
    if (form == FORM_COLUMN) {
        if constexpr (ColCfg<T, NB>::ok) {
            ColGeom g;
            return col_geom<T, NB>(ctx, g);
        }
        return false;
    }
    if (form == FORM_BRICK) return BrickCfg<T, NB>::ok && !ctx->force_coop;
    return true;
}

// nc components of f at every tracer by kernel form `form` (applicable)
template <typename T, int NB>
int sample_form(bfpsh_ctx* ctx, int form, const void* f, int nc, double* d_out) {
    if constexpr (ColCfg<T, NB>::ok) {
        if (form == FORM_COLUMN && (nc == 3 || nc == 9)) {
            ColGeom g;
            col_geom<T, NB>(ctx, g);
            int r = sort_cols<T, NB>(ctx, g);
            if (r) return r;
            if (nc == 3) return launch_col<T, NB, 3, 1>(ctx, g, f, d_out, "sample_col");
            return launch_col<T, NB, 9, 1>(ctx, g, f, d_out, "sample9_col");
        }
    }
    if constexpr (BrickCfg<T, NB>::ok) {
        if (form != FORM_GATHER && !ctx->force_coop) {
            int r = sort_bricks<T, NB>(ctx);
            if (r) return r;
            if constexpr (sizeof(T) == 8)
                if (nc == 1) return launch_brick<T, NB, 1, 1>(ctx, f, d_out, 0.0, 0, "sample1");
            if (nc == 3) return launch_brick<T, NB, 3, 1>(ctx, f, d_out, 0.0, 0, "sample");
            return launch_brick<T, NB, 9, 1>(ctx, f, d_out, 0.0, 0, "sample9");
        }
    }
    {
        int r = sort_tracers(ctx);
        if (r) return r;
    }
    const size_t lds = size_t(2 * NB + 2) * (ctx->deg + 1) * sizeof(double);
    if constexpr (sizeof(T) == 8)
        if (nc == 1) {
            LAUNCH("sample1_gather", (k_sample<T, NB, 1><<<blocks_for(ctx->np, Lanes<NB>::LP), kThreads, lds,
                                                          ctx->stream>>>((const T*)f, ctx->d_pos, d_out,
                                                                         ctx->d_perm + ctx->np, ctx->np, ctx->d_tab,
                                                                         ctx->deg, pgeom_d(ctx))));
            return BFPSH_OK;
        }
    if (nc == 3)
        LAUNCH("sample_gather", (k_sample<T, NB, 3><<<blocks_for(ctx->np, Lanes<NB>::LP), kThreads, lds, ctx->stream>>>(
                                    (const T*)f, ctx->d_pos, d_out, ctx->d_perm + ctx->np, ctx->np, ctx->d_tab, ctx->deg,
                                    pgeom_d(ctx))));
    else
        LAUNCH("sample9_gather", (k_sample<T, NB, 9><<<blocks_for(ctx->np, Lanes<NB>::LP), kThreads, lds,
                                                      ctx->stream>>>((const T*)f, ctx->d_pos, d_out,
                                                                     ctx->d_perm + ctx->np, ctx->np, ctx->d_tab,
                                                                     ctx->deg, pgeom_d(ctx))));
    return BFPSH_OK;
}

// one velocity loop of kernel form `form` (applicable): interpolate the
// 3-plane real field u (ctx->ru in a step) into the head slot of the history
// ring and move the tracers
template <typename T, int NB>
int loop_form(bfpsh_ctx* ctx, int form, const void* u, double dt, int order) {
    if constexpr (ColCfg<T, NB>::ok) {
        if (form == FORM_COLUMN) {
            ColGeom g;
            col_geom<T, NB>(ctx, g);
            int r = sort_cols<T, NB>(ctx, g);
            if (r) return r;
            r = launch_col<T, NB, 3, 0>(ctx, g, u, ctx->d_rhs, "interp_ab_col");
            if (r) return r;
            LAUNCH("ab_moves", (k_ab_moves<<<unsigned((ctx->np + 255) / 256), 256, 0, ctx->stream>>>(
                                   ctx->d_pos, ctx->d_rhs, ctx->np, ctx->nsteps, ctx->head, order, dt)));
            return BFPSH_OK;
        }
    }
    if constexpr (BrickCfg<T, NB>::ok) {
        if (form == FORM_BRICK) {
            int r = sort_bricks<T, NB>(ctx);
            if (r) return r;
            return launch_brick<T, NB, 3, 0>(ctx, u, nullptr, dt, order, "interp_ab");
        }
    }
    if (ctx->sort_age >= ctx->sort_every) {
        int r = sort_tracers(ctx);
        if (r) return r;
    }
    const size_t lds = size_t(2 * NB + 2) * (ctx->deg + 1) * sizeof(double);
    LAUNCH("interp_ab_gather", (k_interp_ab<T, NB><<<blocks_for(ctx->np, Lanes<NB>::LP), kThreads, lds, ctx->stream>>>(
                                   (const T*)u, ctx->d_pos, ctx->d_rhs, ctx->d_perm + ctx->np, ctx->np,
                                   ctx->nsteps, ctx->head, order, ctx->add_head ? 1 : 0, dt, ctx->d_tab, ctx->deg,
                                   pgeom_d(ctx))));
    return BFPSH_OK;
}

// The kernel form of a box's velocity loop and samples: the forced one if it
// applies; else, once per tracer set, the form whose velocity loop (sort,
// interpolation, Adams-Bashforth move of the highest order) runs fastest,
// timed on a copy of the tracer state with dt = 0, when it beats the first
// applicable one by more than 10 % (plan rigor MEASURE or PATIENT, as the
// pass autotuner), or the first applicable of column walk, bricks, global
// gather (ESTIMATE).  The trials interpolate ctx->ru, or the field f of nc
// planes about to be sampled when ctx->ru is not allocated yet; with
// neither (a 1-plane sample first) this call takes the first applicable
// form and the choice waits for a later call.  The forms sum in different
// orders: results agree to rounding, not bit for bit.  BFPSH_TRACER_VERBOSE
// set in the environment prints the trials to stderr.
template <typename T, int NB>
int tracer_form(bfpsh_ctx* ctx, const void* f, int nc, int* form) {
    static const char* const names[4] = {"auto", "bricks", "global gather", "column walk"};
    if (ctx->tracer_kernel != FORM_AUTO && form_ok<T, NB>(ctx, ctx->tracer_kernel)) {
        *form = ctx->tracer_kernel;
        return BFPSH_OK;
    }
    if (ctx->tracer_form != FORM_AUTO) {
        *form = ctx->tracer_form;
        return BFPSH_OK;
    }
    int cand[3], ncand = 0;
    for (int c : {int(FORM_COLUMN), int(FORM_BRICK), int(FORM_GATHER)})
        if (form_ok<T, NB>(ctx, c)) cand[ncand++] = c;
    int best = cand[0];
    const void* u = ctx->ru ? ctx->ru : (nc >= 3 ? f : nullptr);
    if (ncand > 1 && ctx->plan_rigor != BFPSH_ESTIMATE && ctx->tracer_kernel == FORM_AUTO && !u) {
        *form = best;
        return BFPSH_OK;
    }
    if (ncand > 1 && ctx->plan_rigor != BFPSH_ESTIMATE && ctx->tracer_kernel == FORM_AUTO) {
        // the trials move a copy of the positions with a zeroed history; the
        // copy is refreshed before every launch sequence, so that each one
        // locates the tracers from the real positions whatever the previous
        // one left (its dt = 0 move adds 0 x the interpolated values)
        const size_t pos_bytes = size_t(ctx->np) * 3 * sizeof(double);
        const size_t rhs_bytes = size_t(ctx->nsteps) * pos_bytes;
        double *spos = nullptr, *srhs = nullptr;
        hipEvent_t a = nullptr, b = nullptr;
        HIPCHK(hipMalloc((void**)&spos, pos_bytes));
        if (hipMalloc((void**)&srhs, rhs_bytes) != hipSuccess) {
            (void)hipFree(spos);
            return fail(ctx, BFPSH_EHIP, "tracer kernel choice: device memory for the trial state");
        }
        HIPCHK(hipMemsetAsync(srhs, 0, rhs_bytes, ctx->stream));
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        double* const pos = ctx->d_pos;
        double* const rhs = ctx->d_rhs;
        const bool prof = ctx->prof, add_head = ctx->add_head;
        ctx->prof = false;  // the trials stay out of the kernel profile
        ctx->d_pos = spos;
        ctx->d_rhs = srhs;
        ctx->add_head = false;
        auto fresh = [&]() -> int {
            if (hipMemcpyAsync(spos, pos, pos_bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
                return BFPSH_EHIP;
            ctx->sort_age = 1 << 30;  // every form sorts, as in a step
            return BFPSH_OK;
        };
        double tbest = 1e300, tfirst = 0.0;
        int r = BFPSH_OK;
        for (int i = 0; i < ncand && r == BFPSH_OK; i++) {
            r = fresh();
            if (r == BFPSH_OK) r = loop_form<T, NB>(ctx, cand[i], u, 0.0, ctx->nsteps);  // warm-up
            if (r == BFPSH_OK) r = fresh();
            if (r == BFPSH_OK && hipEventRecord(a, ctx->stream) != hipSuccess) r = BFPSH_EHIP;
            if (r == BFPSH_OK) r = loop_form<T, NB>(ctx, cand[i], u, 0.0, ctx->nsteps);
            float ms = 0.f;
            if (r == BFPSH_OK && (hipEventRecord(b, ctx->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                                  hipEventElapsedTime(&ms, a, b) != hipSuccess))
                r = BFPSH_EHIP;
            if (i == 0) tfirst = ms;
            if (r == BFPSH_OK && std::getenv("BFPSH_TRACER_VERBOSE"))
                std::fprintf(stderr, "bfpsh: tracer form trial (n = %d, %s tracers, %lld of them): %s %.3f ms\n",
                             NB, sizeof(T) == 8 ? "FP64" : "FP32", (long long)ctx->np, names[cand[i]], ms);
            if (r == BFPSH_OK && ms < tbest) {
                tbest = ms;
                best = cand[i];
            }
        }
        // another form only when clearly faster: near ties keep the
        // first applicable form, so that runs choose alike
        if (tbest > 0.9 * tfirst) best = cand[0];
        if (r == BFPSH_OK && std::getenv("BFPSH_TRACER_VERBOSE"))
            std::fprintf(stderr, "bfpsh: tracer form chosen: %s\n", names[best]);
        // the trials' launches are done before the copies are freed
        if (r == BFPSH_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) r = BFPSH_EHIP;
        ctx->d_pos = pos;
        ctx->d_rhs = rhs;
        ctx->add_head = add_head;
        ctx->prof = prof;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        (void)hipFree(spos);
        (void)hipFree(srhs);
        if (r) return r == BFPSH_EHIP ? fail(ctx, r, "tracer kernel choice: HIP error") : r;
        ctx->sort_age = 1 << 30;  // the trials' sorts leave another order behind
    }
    ctx->tracer_form = best;
    *form = best;
    return BFPSH_OK;
}

template <typename T, int NB>
int loop_nb(bfpsh_ctx* ctx, double dt, int order) {
    int form = FORM_GATHER;
    int r = tracer_form<T, NB>(ctx, ctx->ru, 3, &form);
    if (r) return r;
    return loop_form<T, NB>(ctx, form, ctx->ru, dt, order);
}

template <typename T, int NB>
int sample_nb(bfpsh_ctx* ctx, const void* f, int nc, double* d_out) {
    int form = FORM_GATHER;
    int r = tracer_form<T, NB>(ctx, f, nc, &form);
    if (r) return r;
    return sample_form<T, NB>(ctx, form, f, nc, d_out);
}

template <typename T>
int loop_t(bfpsh_ctx* ctx, double dt, int order) {
    switch (ctx->pn) {
    case 1: return loop_nb<T, 1>(ctx, dt, order);
    case 2: return loop_nb<T, 2>(ctx, dt, order);
    case 3: return loop_nb<T, 3>(ctx, dt, order);
    case 4: return loop_nb<T, 4>(ctx, dt, order);
    case 5: return loop_nb<T, 5>(ctx, dt, order);
    case 6: return loop_nb<T, 6>(ctx, dt, order);
    case 7: return loop_nb<T, 7>(ctx, dt, order);
    case 8: return loop_nb<T, 8>(ctx, dt, order);
    case 9: return loop_nb<T, 9>(ctx, dt, order);
    case 10: return loop_nb<T, 10>(ctx, dt, order);
    }
    return fail(ctx, BFPSH_EINVAL, "bad neighbour count");
}

template <typename T>
int sample_t(bfpsh_ctx* ctx, const void* f, int nc, double* d_out) {
    switch (ctx->pn) {
    case 1: return sample_nb<T, 1>(ctx, f, nc, d_out);
    case 2: return sample_nb<T, 2>(ctx, f, nc, d_out);
    case 3: return sample_nb<T, 3>(ctx, f, nc, d_out);
    case 4: return sample_nb<T, 4>(ctx, f, nc, d_out);
    case 5: return sample_nb<T, 5>(ctx, f, nc, d_out);
    case 6: return sample_nb<T, 6>(ctx, f, nc, d_out);
    case 7: return sample_nb<T, 7>(ctx, f, nc, d_out);
    case 8: return sample_nb<T, 8>(ctx, f, nc, d_out);
    case 9: return sample_nb<T, 9>(ctx, f, nc, d_out);
    case 10: return sample_nb<T, 10>(ctx, f, nc, d_out);
    }
    return fail(ctx, BFPSH_EINVAL, "bad neighbour count");
}

// The tracer tables of the boxes of one neighbour count: rewrite the
// device table when a box's field, tracers, geometry or time step changed,
// and the device AB state when the host's view of a box (after a put, a
// checkpoint read, or a step taken outside the group) differs from what
// the device holds; due boxes are re-sorted first.
template <int LP>
int prepare_group(const std::vector<bfpsh_ctx*>& boxes, const std::vector<double>& dts, PartBatch& pb) {
    bfpsh_ctx* ctx = boxes[0];  // the shared stream; errors are reported on this box
    const size_t nb = boxes.size();
    std::vector<TracerBox> tb(nb);
    std::vector<int> dyn(4 * nb);
    uint32_t blocks = 0;
    int maxdeg = 0;
    for (size_t i = 0; i < nb; i++) {
        bfpsh_ctx* c = boxes[i];
        if (c->sort_age >= c->sort_every) {
            int r = sort_tracers(c);
            if (r) return fail(ctx, r, c->err);
        }
        TracerBox& b = tb[i];
        std::memset(&b, 0, sizeof(b));
        b.f = c->ru;
        b.pos = c->d_pos;
        b.rhs = c->d_rhs;
        b.perm = c->d_perm + c->np;
        b.tab = c->d_tab;
        b.np = c->np;
        b.pg = pgeom_d(c);
        b.dt = dts[i];
        b.nsteps = c->nsteps;
        b.deg = c->deg;
        b.block0 = blocks;
        blocks += uint32_t((c->np * LP + 255) / 256);
        maxdeg = std::max(maxdeg, c->deg);
        dyn[4 * i + 0] = c->step_idx;
        dyn[4 * i + 1] = c->head;
        dyn[4 * i + 2] = c->add_head ? 1 : 0;
        dyn[4 * i + 3] = 0;
    }
    const size_t bytes = nb * sizeof(TracerBox);
    if (pb.cap < bytes) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        for (void* p : {pb.d, pb.d_dyn})
            if (p) HIPCHK(hipFree(p));
        pb.d = pb.d_dyn = nullptr;
        HIPCHK(hipMalloc(&pb.d, bytes));
        HIPCHK(hipMalloc(&pb.d_dyn, nb * sizeof(TracerDyn)));
        pb.cap = bytes;
        pb.tab.clear();
        pb.dyn.clear();
    }
    const unsigned char* tbb = reinterpret_cast<const unsigned char*>(tb.data());
    if (pb.tab.size() != bytes || std::memcmp(pb.tab.data(), tbb, bytes) != 0) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipMemcpy(pb.d, tb.data(), bytes, hipMemcpyHostToDevice));
        pb.tab.assign(tbb, tbb + bytes);
    }
    if (pb.dyn != dyn) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipMemcpy(pb.d_dyn, dyn.data(), dyn.size() * sizeof(int), hipMemcpyHostToDevice));
        pb.dyn = dyn;
    }
    pb.nboxes = int(nb);
    pb.blocks = blocks;
    pb.maxdeg = maxdeg;
    return BFPSH_OK;
}

template <typename T, int NB>
int prepare_nb(const std::vector<bfpsh_ctx*>& boxes, const std::vector<double>& dts, PartBatch& pb) {
    int r = prepare_group<Lanes<NB>::LP>(boxes, dts, pb);
    if (r) return r;
    pb.lds = size_t(2 * NB + 2) * (pb.maxdeg + 1) * sizeof(double);
    return BFPSH_OK;
}

template <typename T, int NB>
int launch_nb(bfpsh_ctx* ctx, PartBatch& pb) {
    const TracerBox* tb = static_cast<const TracerBox*>(pb.d);
    TracerDyn* dyn = static_cast<TracerDyn*>(pb.d_dyn);
    LAUNCH("interp_ab_multi", (k_interp_ab_multi<T, NB><<<pb.blocks, kThreads, pb.lds, ctx->stream>>>(
                                  tb, dyn, pb.nboxes)));
    LAUNCH("tracer_advance", (k_tracer_dyn_advance<<<(pb.nboxes + 255) / 256, 256, 0, ctx->stream>>>(
                                 tb, dyn, pb.nboxes)));
    return BFPSH_OK;
}

#define BFPSH_NB_SWITCH(CALL)                                  \
    switch (nbr) {                                             \
    case 1: return CALL(1); case 2: return CALL(2);            \
    case 3: return CALL(3); case 4: return CALL(4);            \
    case 5: return CALL(5); case 6: return CALL(6);            \
    case 7: return CALL(7); case 8: return CALL(8);            \
    case 9: return CALL(9); case 10: return CALL(10);          \
    }

template <typename T>
int prepare_t(const std::vector<bfpsh_ctx*>& boxes, const std::vector<double>& dts, PartBatch& pb, int nbr) {
#define BFPSH_P(K) prepare_nb<T, K>(boxes, dts, pb)
    BFPSH_NB_SWITCH(BFPSH_P)
#undef BFPSH_P
    return fail(boxes[0], BFPSH_EINVAL, "bad neighbour count");
}

template <typename T>
int launch_t(bfpsh_ctx* ctx, PartBatch& pb, int nbr) {
#define BFPSH_L(K) launch_nb<T, K>(ctx, pb)
    BFPSH_NB_SWITCH(BFPSH_L)
#undef BFPSH_L
    return fail(ctx, BFPSH_EINVAL, "bad neighbour count");
}
#undef BFPSH_NB_SWITCH

}  // namespace

int bfpsh_part_ens_prepare(const std::vector<bfpsh_ctx*>& boxes, const std::vector<double>& dts, PartBatch* pb,
                           int nbatch, int* ngroups) {
    // one group (one table, one launch) per neighbour count, in order of
    // first appearance
    std::vector<int> counts;
    for (bfpsh_ctx* c : boxes)
        if (std::find(counts.begin(), counts.end(), c->pn) == counts.end()) counts.push_back(c->pn);
    if (int(counts.size()) > nbatch) return fail(boxes[0], BFPSH_EINVAL, "too many neighbour counts in one ensemble");
    for (size_t gi = 0; gi < counts.size(); gi++) {
        std::vector<bfpsh_ctx*> grp;
        std::vector<double> gdt;
        for (size_t i = 0; i < boxes.size(); i++)
            if (boxes[i]->pn == counts[gi]) {
                grp.push_back(boxes[i]);
                gdt.push_back(dts[i]);
            }
        pb[gi].nbr = counts[gi];
        pb[gi].prec = grp[0]->prec;
        pb[gi].ctx = grp[0];
        int r = grp[0]->prec == BFPSH_FP32 ? prepare_t<float>(grp, gdt, pb[gi], counts[gi])
                                           : prepare_t<double>(grp, gdt, pb[gi], counts[gi]);
        if (r) return r;
    }
    *ngroups = int(counts.size());
    return BFPSH_OK;
}

int bfpsh_part_ens_launch(PartBatch* pb, int ngroups) {
    for (int gi = 0; gi < ngroups; gi++) {
        bfpsh_ctx* ctx = pb[gi].ctx;
        int r = pb[gi].prec == BFPSH_FP32 ? launch_t<float>(ctx, pb[gi], pb[gi].nbr)
                                          : launch_t<double>(ctx, pb[gi], pb[gi].nbr);
        if (r) return r;
    }
    return BFPSH_OK;
}

void bfpsh_part_ens_advanced(const std::vector<bfpsh_ctx*>& boxes, PartBatch* pb, int ngroups) {
    // the host's view of every box follows the device's k_tracer_dyn_advance
    for (bfpsh_ctx* c : boxes) {
        c->sort_age += 1;
        c->step_idx += 1;
        c->head = (c->head + c->nsteps - 1) % c->nsteps;
        c->head_zero = true;
        c->add_head = false;
    }
    for (int gi = 0; gi < ngroups; gi++) {
        PartBatch& g = pb[gi];
        const TracerBox* tb = reinterpret_cast<const TracerBox*>(g.tab.data());
        for (int i = 0; i < g.nboxes && 4 * i + 3 < int(g.dyn.size()); i++) {
            const int ns = tb[i].nsteps;
            g.dyn[4 * i + 0] += 1;
            g.dyn[4 * i + 1] = (g.dyn[4 * i + 1] + ns - 1) % ns;
            g.dyn[4 * i + 2] = 0;
        }
    }
}

void bfpsh_part_batch_free(PartBatch& pb) {
    for (void* p : {pb.d, pb.d_dyn})
        if (p) (void)hipFree(p);
    pb.d = pb.d_dyn = nullptr;
    pb.cap = 0;
    pb.tab.clear();
    pb.dyn.clear();
}

int bfpsh_part_release(bfpsh_ctx* ctx) {
    for (void* p : {(void*)ctx->d_keys, (void*)ctx->d_perm, ctx->d_sort_tmp, (void*)ctx->d_bstart, (void*)ctx->d_crec})
        if (p) (void)hipFree(p);
    ctx->d_crec = nullptr;
    ctx->d_bstart = nullptr;
    ctx->nbricks = 0;
    ctx->d_keys = nullptr;
    ctx->d_perm = nullptr;
    ctx->d_sort_tmp = nullptr;
    ctx->sort_tmp_bytes = 0;
    ctx->sort_age = 1 << 30;
    return BFPSH_OK;
}

// completeLoop: compute (interpolate into the head of the history ring) +
// move (Adams-Bashforth of order min(step_idx, nsteps)) + inc_step_idx +
// shift_rhs_vectors (the oldest slot becomes the new, logically zero, head)
int bfpsh_part_loop(bfpsh_ctx* ctx, double dt) {
    if (ctx->np == 0) return BFPSH_OK;
    const int order = std::min(ctx->step_idx, ctx->nsteps);
    int r = ctx->prec == BFPSH_FP32 ? loop_t<float>(ctx, dt, order) : loop_t<double>(ctx, dt, order);
    if (r) return r;
    ctx->sort_age += 1;
    ctx->step_idx += 1;
    ctx->head = (ctx->head + ctx->nsteps - 1) % ctx->nsteps;
    ctx->head_zero = true;
    ctx->add_head = false;
    return BFPSH_OK;
}

int bfpsh_part_sample(bfpsh_ctx* ctx, const void* f, int nc, double* d_out) {
    if (ctx->np == 0) return BFPSH_OK;
    if (nc != 3 && nc != 9 && !(nc == 1 && ctx->prec == BFPSH_FP64))
        return fail(ctx, BFPSH_EINVAL, "sampling needs 3 or 9 components (or 1 in FP64)");
    return ctx->prec == BFPSH_FP32 ? sample_t<float>(ctx, f, nc, d_out) : sample_t<double>(ctx, f, nc, d_out);
}
